// Metadata filters for where=: the filter expressions of the reference
// (include/vecengine/metadata.hpp:52-162, Expr; SQL WHERE parsing,
// src/metadata.cpp:126-403) and their evaluation into a slot bitmap over an
// index's metadata columns (MetadataStore::evaluate, src/metadata.cpp:974-1487),
// plus the checks and appends of add(metadata=) (MetadataStore::check /
// append, src/metadata.cpp:515-800). Host code: the bitmap it makes is the
// allowed-slot filter the device search takes.
#pragma once

#include "qfile.hpp"

#include <cstdint>
#include <memory>
#include <string>
#include <string_view>
#include <unordered_map>
#include <utility>
#include <variant>
#include <vector>

namespace qh {

// One value of a comparison: an int, a float or a string (the reference's
// Expr::Value, whose order matches the column types 0, 1, 2).
using FilterValue = std::variant<int64_t, double, std::string>;

class Expr {
public:
    enum class Op : uint8_t { Eq, Ne, Lt, Le, Gt, Ge, In, IsNull, And, Or, Not };
    // deepest nesting a filter may have (Expr::kMaxDepth); And / Or chains are flat
    static constexpr size_t kMaxDepth = 256;
    // conditions a parsed SQL string may hold by default (Expr::kDefaultMaxConditions)
    static constexpr size_t kDefaultMaxConditions = 1024;

    static Expr compare(std::string field, Op op, FilterValue value);
    static Expr isin(std::string field, std::vector<FilterValue> values);
    static Expr is_null(std::string field);
    // (field >= lo) & (field <= hi)
    static Expr between(std::string field, FilterValue lo, FilterValue hi);
    // all (And) or any (Or) of the filters as one node, same-op children merged in
    static Expr join(Op op, std::vector<Expr> children);
    static Expr negate(const Expr& e);
    // A SQL WHERE clause with ? (list) or :name (named) placeholders; at most
    // one of list / named is given. Throws std::invalid_argument naming the
    // column of a syntax error.
    static Expr parse(std::string_view sql, const std::vector<FilterValue>* list,
                      const std::unordered_map<std::string, FilterValue>* named);
    static size_t max_conditions();
    static void set_max_conditions(size_t n);

    Op op() const { return node_->op; }
    const std::string& field() const { return node_->field; }
    const std::vector<FilterValue>& values() const { return node_->values; }
    const std::vector<Expr>& children() const { return node_->children; }
    size_t depth() const { return node_->depth; }
    // the readable form qann prints, e.g. ((price < 50) & ~is_null(color))
    std::string to_string() const;

private:
    struct Node {
        Op op;
        std::string field;
        std::vector<FilterValue> values;
        std::vector<Expr> children;
        size_t depth = 1;
    };
    explicit Expr(std::shared_ptr<const Node> n) : node_(std::move(n)) {}
    std::shared_ptr<const Node> node_;
};

// The bitmap (ceil(slots / 64) words) of slots whose metadata rows match e.
// Deleted slots are not cleared. Throws std::invalid_argument for a column the
// store does not have or a value of the wrong kind.
std::vector<uint64_t> evaluate(const Metadata& md, const Expr& e, size_t slots);

// One column of an add(metadata=): its values for the add's n rows, and
// valid (empty: every row has a value; else n flags, 0 for null). type is
// the kind of the values given (0 int, 1 float, 2 string).
struct MetaInput {
    std::string name;
    uint8_t type = 0;
    std::vector<int64_t> ints;
    std::vector<double> floats;
    std::vector<std::string> strings;
    std::vector<uint8_t> valid;
    size_t size() const { return type == 0 ? ints.size() : type == 1 ? floats.size() : strings.size(); }
};

// MetadataStore::check: n values per column, unique non-empty names, a column
// that exists keeps its type (ints may go into a float column), a new column
// has a value to take its type from. Throws std::invalid_argument.
void check_meta(const Metadata& md, size_t n, const std::vector<MetaInput>& in);
// MetadataStore::append: writes the add's n rows at rows .. rows + n - 1 of
// columns that already hold them as nulls (the add appended nulls to every
// column); columns new to the store are created first (in the order given),
// null for every row. Columns the add leaves out keep their nulls. An add of
// no rows creates no columns. check_meta first.
void append_meta(Metadata& md, size_t rows, size_t n, const std::vector<MetaInput>& in);

// Each column's name and type name ("int", "float", "str"), in creation order.
std::vector<std::pair<std::string, std::string>> meta_schema(const Metadata& md);

} // namespace qh
