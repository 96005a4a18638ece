"""Shared descriptions of the golden fixtures for the test modules."""
import json

import numpy as np

from conftest import REF_DATA, SYNTH

# The reference's own saved indexes (tests/data of maxtmiller/qann) with the
# queries and results it recorded: (file, queries, ids must match exactly).
REF_FIXTURES = [
    ("v4_ivf_fs", "v4_queries"),
    ("v11_ivf_cos_fs", "v11_cos_queries"),
    ("v12_ivf_cos_l2", "v12_cos_queries"),
    ("v9_ip_refine_ivf_fs", "v9_ip_queries"),
    ("v1_flat", "v1_queries"),
    ("v2_flat", "v2_queries"),
    ("v3_flat", "v3_queries"),
    ("v4_flat", "v4_queries"),
    ("v5_flat", "v5_queries"),
    ("v9_ip_flat", "v9_ip_queries"),
    ("v10_refine", "v10_queries"),
]
# files of other IVF storage kinds (8-bit PQ, SQ, float lists) and their refines
REF_FIXTURES_OTHER = [
    ("v1_ivf_pq", "v1_queries"),
    ("v2_ivf_pq", "v2_queries"),
    ("v3_ivf_pq", "v3_queries"),
    ("v1_refine", "v1_queries"),
    ("v2_refine", "v2_queries"),
    ("v3_refine", "v3_queries"),
    ("v8_ivf_rot", "v8_queries"),
    ("v7_ivf_sq8", "v7_queries"),
    ("v7_ivf_sq4", "v7_queries"),
    ("v11_ivf_sq8", "v11_queries"),
    ("v11_ivf_float", "v11_queries"),
]

SYNTH_MANIFEST = json.loads((SYNTH / "manifest.json").read_text())
SYNTH_CASES = sorted(SYNTH_MANIFEST)


def ref_fixture(name, queries):
    q = np.load(REF_DATA / f"{queries}.npy")
    ids = np.load(REF_DATA / f"{name}_ids.npy")
    dists = np.load(REF_DATA / f"{name}_dists.npy")
    return REF_DATA / f"{name}.qann", q, ids, dists


def synth_case(name):
    """(path, queries, expected dict, parameter dict) of a golden case."""
    q = np.load(SYNTH / f"{name}_queries.npy")
    with np.load(SYNTH / f"{name}_expected.npz", allow_pickle=False) as z:
        exp = {k: z[k] for k in z.files}
    return SYNTH / f"{name}.qann", q, exp, SYNTH_MANIFEST[name]


def synth_queries(exp, params):
    """Yields (nprobe or None, k, filter name, allowed ids or None, key) for every recorded result."""
    for key in exp:
        if not key.endswith("_ids") or key.startswith("filter_"):
            continue
        stem = key[: -len("_ids")]
        p, k, f = stem.split("_", 2)
        npb = int(p[1:]) or None
        allowed = None if f == "none" else exp[f"filter_{f}"]
        yield npb, int(k[1:]), f, allowed, stem
