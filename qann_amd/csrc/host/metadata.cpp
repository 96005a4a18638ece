// Metadata filters (metadata.hpp).
#include "metadata.hpp"

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <unordered_set>

namespace qh {

namespace {

std::atomic<size_t> g_max_conditions{Expr::kDefaultMaxConditions};

bool is_comparison(Expr::Op op) { return op <= Expr::Op::Ge; }

void check_depth(size_t depth) {
    if (depth > Expr::kMaxDepth)
        throw std::invalid_argument("filter nested more than " + std::to_string(Expr::kMaxDepth) + " levels deep");
}

// A value as qann prints it: ints in decimal, floats as an ostream's default
// (six significant digits) with ".0" after whole numbers below 1e15, strings
// in single quotes.
std::string show(const FilterValue& v) {
    if (const auto* i = std::get_if<int64_t>(&v)) return std::to_string(*i);
    if (const auto* d = std::get_if<double>(&v)) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", *d);
        std::string s(buf);
        if (std::isfinite(*d) && *d == std::floor(*d) && std::fabs(*d) < 1e15) s += ".0";
        return s;
    }
    return "'" + std::get<std::string>(v) + "'";
}

const char* op_text(Expr::Op op) {
    switch (op) {
    case Expr::Op::Eq: return "==";
    case Expr::Op::Ne: return "!=";
    case Expr::Op::Lt: return "<";
    case Expr::Op::Le: return "<=";
    case Expr::Op::Gt: return ">";
    case Expr::Op::Ge: return ">=";
    default: return "?";
    }
}

// The comparison that holds exactly where op fails on values that compare
// (SQL's NOT of a comparison; nulls fail both).
Expr::Op flipped(Expr::Op op) {
    switch (op) {
    case Expr::Op::Eq: return Expr::Op::Ne;
    case Expr::Op::Ne: return Expr::Op::Eq;
    case Expr::Op::Lt: return Expr::Op::Ge;
    case Expr::Op::Le: return Expr::Op::Gt;
    case Expr::Op::Gt: return Expr::Op::Le;
    default: return Expr::Op::Lt;
    }
}

// Recursive descent over the SQL WHERE grammar of Expr::parse:
//   or   := and (OR and)*          and := unary (AND unary)*
//   unary := NOT unary | '(' or ')' | predicate
//   predicate := name IS [NOT] NULL | name [NOT] IN '(' value, ... ')'
//              | name [NOT] BETWEEN value AND value | name cmp value
// NOT is carried down as a flag (negated) instead of building ~ nodes: a
// negated comparison flips its operator, negated IN / BETWEEN exclude nulls,
// negated IS NULL is ~IS NULL, and AND / OR swap (De Morgan) -- SQL's
// three-valued NOT, under which a null row matches neither a condition nor its
// negation.
class Sql {
public:
    Sql(std::string_view text, const std::vector<FilterValue>* list,
        const std::unordered_map<std::string, FilterValue>* named)
        : t_(text), list_(list), named_(named), limit_(Expr::max_conditions()) {}

    Expr run() {
        Expr e = disjunction(false);
        blanks();
        if (at_ < t_.size()) error("unexpected '" + next_token() + "'");
        if (list_ && taken_ != list_->size())
            throw std::invalid_argument("filter has " + std::to_string(taken_) + " ? placeholders but " +
                                        std::to_string(list_->size()) + " parameters were given");
        return e;
    }

private:
    [[noreturn]] void error(const std::string& what) const {
        throw std::invalid_argument(what + (at_ < t_.size() ? " at column " + std::to_string(at_ + 1) : " at end of filter"));
    }
    static bool letter(char c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); }
    static bool digit(char c) { return c >= '0' && c <= '9'; }
    static bool in_word(char c) { return letter(c) || digit(c) || c == '_' || c == '.'; }
    static char up(char c) { return c >= 'a' && c <= 'z' ? static_cast<char>(c - 32) : c; }

    void blanks() {
        while (at_ < t_.size() && (t_[at_] == ' ' || t_[at_] == '\t' || t_[at_] == '\n' || t_[at_] == '\r' ||
                                   t_[at_] == '\f' || t_[at_] == '\v'))
            ++at_;
    }
    // the word (or the one character, whole if UTF-8) at the cursor, for messages
    std::string next_token() {
        blanks();
        size_t e = at_;
        while (e < t_.size() && in_word(t_[e])) ++e;
        if (e == at_ && at_ < t_.size()) {
            ++e;
            while (e < t_.size() && (static_cast<unsigned char>(t_[e]) & 0xC0) == 0x80) ++e;
        }
        return std::string(t_.substr(at_, e - at_));
    }
    // a keyword, case-insensitive and whole-word
    bool word(std::string_view kw) {
        blanks();
        if (t_.size() - at_ < kw.size()) return false;
        for (size_t i = 0; i < kw.size(); ++i)
            if (up(t_[at_ + i]) != kw[i]) return false;
        if (at_ + kw.size() < t_.size() && in_word(t_[at_ + kw.size()])) return false;
        at_ += kw.size();
        return true;
    }
    bool punct(std::string_view p) {
        blanks();
        if (t_.substr(at_, p.size()) != p) return false;
        at_ += p.size();
        return true;
    }
    static bool reserved(const std::string& w) {
        std::string u(w);
        for (char& c : u) c = up(c);
        static const char* kws[] = {"AND", "OR", "NOT", "IN", "IS", "NULL", "BETWEEN"};
        for (const char* k : kws)
            if (u == k) return true;
        return false;
    }

    Expr disjunction(bool negated) {
        std::vector<Expr> parts{conjunction(negated)};
        while (word("OR")) parts.push_back(conjunction(negated));
        return Expr::join(negated ? Expr::Op::And : Expr::Op::Or, std::move(parts));
    }
    Expr conjunction(bool negated) {
        std::vector<Expr> parts{nested(negated)};
        while (word("AND")) parts.push_back(nested(negated));
        return Expr::join(negated ? Expr::Op::Or : Expr::Op::And, std::move(parts));
    }
    // NOT and parentheses recurse: their nesting is bounded here, with a column
    Expr nested(bool negated) {
        if (++depth_ > Expr::kMaxDepth)
            error("filter nested more than " + std::to_string(Expr::kMaxDepth) + " levels deep");
        Expr e = unary(negated);
        --depth_;
        return e;
    }
    Expr unary(bool negated) {
        if (word("NOT")) return nested(!negated);
        if (punct("(")) {
            Expr e = disjunction(negated);
            if (!punct(")")) error("expected ')'");
            return e;
        }
        return predicate(negated);
    }

    std::string quoted(char q) {
        const size_t open = at_++;
        std::string out;
        for (;;) {
            if (at_ >= t_.size()) {
                at_ = open;
                error(q == '\'' ? "unterminated string" : "unterminated quoted name");
            }
            const char c = t_[at_];
            if (c == q) {
                if (at_ + 1 < t_.size() && t_[at_ + 1] == q) {
                    out += q;
                    at_ += 2;
                    continue;
                }
                ++at_;
                return out;
            }
            out += c;
            ++at_;
        }
    }
    std::string column() {
        blanks();
        if (at_ < t_.size() && t_[at_] == '"') return quoted('"');
        const size_t start = at_;
        if (at_ < t_.size() && (letter(t_[at_]) || t_[at_] == '_'))
            while (at_ < t_.size() && in_word(t_[at_])) ++at_;
        if (at_ < t_.size() && static_cast<unsigned char>(t_[at_]) >= 0x80) {
            at_ = start;
            error("a column name with non-ASCII characters needs double quotes");
        }
        std::string name(t_.substr(start, at_ - start));
        if (name.empty() || reserved(name)) {
            at_ = start;
            error("expected a column name");
        }
        return name;
    }
    FilterValue value() {
        blanks();
        if (at_ >= t_.size()) error("expected a value");
        const char c = t_[at_];
        if (c == '\'') return quoted('\'');
        if (c == '?') {
            if (!list_) error(named_ ? "? placeholder, but parameters were given by name"
                                     : "? placeholder, but no parameters were given");
            if (taken_ >= list_->size()) error("more ? placeholders than parameters");
            ++at_;
            return (*list_)[taken_++];
        }
        if (c == ':') {
            size_t e = at_ + 1;
            while (e < t_.size() && in_word(t_[e])) ++e;
            const std::string key(t_.substr(at_ + 1, e - at_ - 1));
            if (key.empty()) error("expected a parameter name after ':'");
            if (!named_) error(list_ ? ":" + key + " placeholder, but parameters were given as a list"
                                     : ":" + key + " placeholder, but no parameters were given");
            const auto it = named_->find(key);
            if (it == named_->end()) error("no parameter named '" + key + "'");
            at_ = e;
            return it->second;
        }
        // sign, digits, an optional fraction, an optional exponent
        const size_t start = at_;
        size_t i = at_;
        if (t_[i] == '+' || t_[i] == '-') ++i;
        const size_t lead = i;
        bool frac = false;
        while (i < t_.size() && digit(t_[i])) ++i;
        if (i < t_.size() && t_[i] == '.') {
            frac = true;
            ++i;
            while (i < t_.size() && digit(t_[i])) ++i;
        }
        if (i == lead || (frac && i == lead + 1)) error("expected a value");
        if (i < t_.size() && (t_[i] == 'e' || t_[i] == 'E')) {
            size_t j = i + 1;
            if (j < t_.size() && (t_[j] == '+' || t_[j] == '-')) ++j;
            if (j < t_.size() && digit(t_[j])) {
                frac = true;
                for (i = j; i < t_.size() && digit(t_[i]); ++i) {
                }
            }
        }
        if (i < t_.size() && in_word(t_[i])) error("expected a value");
        const std::string text(t_.substr(start, i - start));
        at_ = i;
        if (frac) {
            const double d = std::strtod(text.c_str(), nullptr);
            if (std::isinf(d)) {
                at_ = start;
                error("number out of range");
            }
            return d;
        }
        errno = 0;
        const long long v = std::strtoll(text.c_str(), nullptr, 10);
        if (errno == ERANGE) {
            at_ = start;
            error("integer out of range");
        }
        return static_cast<int64_t>(v);
    }
    Expr predicate(bool negated) {
        if (++conditions_ > limit_)
            error("filter has more than " + std::to_string(limit_) +
                  " conditions; combine values with IN (...) or raise the limit");
        const std::string name = column();
        if (word("IS")) {
            const bool is_not = word("NOT");
            if (!word("NULL")) error("expected NULL");
            const Expr e = Expr::is_null(name);
            return negated != is_not ? Expr::negate(e) : e;
        }
        const bool not_kw = word("NOT");
        if (not_kw) negated = !negated;
        if (word("IN")) {
            if (!punct("(")) error("expected '('");
            std::vector<FilterValue> vs{value()};
            while (punct(",")) vs.push_back(value());
            if (!punct(")")) error("expected ',' or ')'");
            const Expr e = Expr::isin(name, std::move(vs));
            return negated ? Expr::join(Expr::Op::And, {Expr::negate(e), Expr::negate(Expr::is_null(name))}) : e;
        }
        if (word("BETWEEN")) {
            FilterValue lo = value();
            if (!word("AND")) error("expected AND");
            FilterValue hi = value();
            if (negated)
                return Expr::join(Expr::Op::Or, {Expr::compare(name, Expr::Op::Lt, std::move(lo)),
                                                 Expr::compare(name, Expr::Op::Gt, std::move(hi))});
            return Expr::between(name, std::move(lo), std::move(hi));
        }
        if (not_kw) error("expected IN or BETWEEN");
        Expr::Op op;
        if (punct("<=")) op = Expr::Op::Le;
        else if (punct(">=")) op = Expr::Op::Ge;
        else if (punct("<>") || punct("!=")) op = Expr::Op::Ne;
        else if (punct("==") || punct("=")) op = Expr::Op::Eq;
        else if (punct("<")) op = Expr::Op::Lt;
        else if (punct(">")) op = Expr::Op::Gt;
        else error("expected a comparison, IN, BETWEEN or IS");
        return Expr::compare(name, negated ? flipped(op) : op, value());
    }

    std::string_view t_;
    size_t at_ = 0;
    const std::vector<FilterValue>* list_;
    const std::unordered_map<std::string, FilterValue>* named_;
    size_t taken_ = 0, depth_ = 0, conditions_ = 0, limit_;
};

} // namespace

Expr Expr::compare(std::string field, Op op, FilterValue value) {
    if (!is_comparison(op)) throw std::invalid_argument("compare needs ==, !=, <, <=, > or >=");
    return Expr(std::make_shared<const Node>(Node{op, std::move(field), {std::move(value)}, {}, 1}));
}

Expr Expr::isin(std::string field, std::vector<FilterValue> values) {
    return Expr(std::make_shared<const Node>(Node{Op::In, std::move(field), std::move(values), {}, 1}));
}

Expr Expr::is_null(std::string field) {
    return Expr(std::make_shared<const Node>(Node{Op::IsNull, std::move(field), {}, {}, 1}));
}

Expr Expr::between(std::string field, FilterValue lo, FilterValue hi) {
    Expr ge = compare(field, Op::Ge, std::move(lo));
    return join(Op::And, {std::move(ge), compare(std::move(field), Op::Le, std::move(hi))});
}

Expr Expr::join(Op op, std::vector<Expr> children) {
    if (op != Op::And && op != Op::Or) throw std::invalid_argument("join needs And or Or");
    if (children.empty()) throw std::invalid_argument("join needs at least one filter");
    if (children.size() == 1) return std::move(children.front());
    std::vector<Expr> flat;
    size_t below = 0;
    for (Expr& c : children) {
        if (c.op() == op) {
            // a child of the same op: its own children join this node directly
            below = std::max(below, c.depth() - 1);
            flat.insert(flat.end(), c.children().begin(), c.children().end());
        } else {
            below = std::max(below, c.depth());
            flat.push_back(std::move(c));
        }
    }
    check_depth(below + 1);
    return Expr(std::make_shared<const Node>(Node{op, {}, {}, std::move(flat), below + 1}));
}

Expr Expr::negate(const Expr& e) {
    check_depth(e.depth() + 1);
    return Expr(std::make_shared<const Node>(Node{Op::Not, {}, {}, {e}, e.depth() + 1}));
}

Expr Expr::parse(std::string_view sql, const std::vector<FilterValue>* list,
                 const std::unordered_map<std::string, FilterValue>* named) {
    if (list && list->empty()) list = nullptr;
    return Sql(sql, list, named).run();
}

size_t Expr::max_conditions() { return g_max_conditions.load(); }

void Expr::set_max_conditions(size_t n) {
    if (n == 0) throw std::invalid_argument("the filter condition limit must be at least 1");
    g_max_conditions.store(n);
}

std::string Expr::to_string() const {
    switch (op()) {
    case Op::And:
    case Op::Or: {
        std::string s = "(";
        for (size_t i = 0; i < children().size(); ++i) {
            if (i) s += op() == Op::And ? " & " : " | ";
            s += children()[i].to_string();
        }
        return s + ")";
    }
    case Op::Not: return "~" + children()[0].to_string();
    case Op::IsNull: return "is_null(" + field() + ")";
    case Op::In: {
        std::string s = field() + " in [";
        for (size_t i = 0; i < values().size(); ++i) s += (i ? ", " : "") + show(values()[i]);
        return s + "]";
    }
    default: return "(" + field() + " " + op_text(op()) + " " + show(values()[0]) + ")";
    }
}

// ---------------------------------------------------------------------------
// evaluation
// ---------------------------------------------------------------------------

namespace {

using Words = std::vector<uint64_t>;

template <class T>
bool holds(Expr::Op op, const T& a, const T& b) {
    switch (op) {
    case Expr::Op::Eq: return a == b;
    case Expr::Op::Ne: return a != b;
    case Expr::Op::Lt: return a < b;
    case Expr::Op::Le: return a <= b;
    case Expr::Op::Gt: return a > b;
    case Expr::Op::Ge: return a >= b;
    default: return false;
    }
}

double number(const FilterValue& v, const std::string& field) {
    if (const auto* i = std::get_if<int64_t>(&v)) return static_cast<double>(*i);
    if (const auto* d = std::get_if<double>(&v)) return *d;
    throw std::invalid_argument("metadata field '" + field + "' holds numbers; got the string " + show(v));
}

const std::string& text(const FilterValue& v, const std::string& field) {
    if (const auto* s = std::get_if<std::string>(&v)) return *s;
    throw std::invalid_argument("metadata field '" + field + "' holds strings; got " + show(v));
}

// words with bit i set where valid[i] and pass(i), for i < slots
template <class Pass>
Words rows_where(const MetaColumn& col, size_t slots, Pass&& pass) {
    Words w((slots + 63) / 64, 0);
    for (size_t i = 0; i < slots; ++i)
        if (col.valid[i] && pass(i)) w[i >> 6] |= uint64_t{1} << (i & 63);
    return w;
}

Words leaf(const Metadata& md, const Expr& e, size_t slots) {
    const MetaColumn* col = nullptr;
    for (const MetaColumn& c : md.columns)
        if (c.name == e.field()) col = &c;
    if (!col) throw std::invalid_argument("unknown metadata field '" + e.field() + "'");
    if (col->valid.size() < slots) throw std::logic_error("metadata column shorter than the index");
    const Expr::Op op = e.op();
    if (op == Expr::Op::IsNull) {
        Words w((slots + 63) / 64, 0);
        for (size_t i = 0; i < slots; ++i)
            if (!col->valid[i]) w[i >> 6] |= uint64_t{1} << (i & 63);
        return w;
    }
    if (col->type == 2) {
        // strings: decide each dictionary entry once, then look rows up by code
        std::vector<uint8_t> hit(col->dict.size(), 0);
        if (op == Expr::Op::In || op == Expr::Op::Eq) {
            std::unordered_set<std::string> wanted;
            for (const FilterValue& v : e.values()) wanted.insert(text(v, e.field()));
            for (size_t c = 0; c < col->dict.size(); ++c) hit[c] = wanted.count(col->dict[c]) ? 1 : 0;
        } else {
            const std::string& v = text(e.values()[0], e.field());
            for (size_t c = 0; c < col->dict.size(); ++c) hit[c] = holds(op, col->dict[c], v) ? 1 : 0;
        }
        return rows_where(*col, slots, [&](size_t i) { return col->codes[i] < hit.size() && hit[col->codes[i]]; });
    }
    if (col->type == 0) {
        const std::vector<int64_t>& x = col->ints;
        if (op == Expr::Op::In) {
            // ints match exactly; a float value matches an int equal to it as a double
            std::unordered_set<int64_t> ints;
            std::vector<double> floats;
            for (const FilterValue& v : e.values()) {
                if (const auto* i = std::get_if<int64_t>(&v)) ints.insert(*i);
                else floats.push_back(number(v, e.field()));
            }
            return rows_where(*col, slots, [&](size_t i) {
                if (ints.count(x[i])) return true;
                const double xd = static_cast<double>(x[i]);
                for (double f : floats)
                    if (xd == f) return true;
                return false;
            });
        }
        if (const auto* iv = std::get_if<int64_t>(&e.values()[0])) {
            const int64_t v = *iv;
            return rows_where(*col, slots, [&](size_t i) { return holds(op, x[i], v); });
        }
        const double v = number(e.values()[0], e.field());
        return rows_where(*col, slots, [&](size_t i) { return holds(op, static_cast<double>(x[i]), v); });
    }
    const std::vector<double>& x = col->floats;
    if (op == Expr::Op::In) {
        std::vector<double> wanted;
        for (const FilterValue& v : e.values()) wanted.push_back(number(v, e.field()));
        return rows_where(*col, slots, [&](size_t i) {
            for (double f : wanted)
                if (x[i] == f) return true;
            return false;
        });
    }
    const double v = number(e.values()[0], e.field());
    return rows_where(*col, slots, [&](size_t i) { return holds(op, x[i], v); });
}

Words tree(const Metadata& md, const Expr& e, size_t slots) {
    const size_t nw = (slots + 63) / 64;
    switch (e.op()) {
    case Expr::Op::And:
    case Expr::Op::Or: {
        const bool all = e.op() == Expr::Op::And;
        Words acc = tree(md, e.children()[0], slots);
        for (size_t c = 1; c < e.children().size(); ++c) {
            const Words w = tree(md, e.children()[c], slots);
            for (size_t i = 0; i < nw; ++i) acc[i] = all ? acc[i] & w[i] : acc[i] | w[i];
        }
        return acc;
    }
    case Expr::Op::Not: {
        Words w = tree(md, e.children()[0], slots);
        for (size_t i = 0; i < nw; ++i) w[i] = ~w[i];
        if (slots % 64) w.back() &= (uint64_t{1} << (slots % 64)) - 1;
        return w;
    }
    default: return leaf(md, e, slots);
    }
}

} // namespace

std::vector<uint64_t> evaluate(const Metadata& md, const Expr& e, size_t slots) { return tree(md, e, slots); }

// ---------------------------------------------------------------------------
// add(metadata=)
// ---------------------------------------------------------------------------

namespace {

const char* kind_name(uint8_t t) { return t == 0 ? "int" : t == 1 ? "float" : "string"; }

// ints fit a float column (converted); nothing else crosses kinds
bool fits(uint8_t column, uint8_t given) { return column == given || (column == 1 && given == 0); }

} // namespace

void check_meta(const Metadata& md, size_t n, const std::vector<MetaInput>& in) {
    std::unordered_set<std::string> names;
    for (const MetaInput& c : in) {
        if (c.name.empty()) throw std::invalid_argument("metadata column names must not be empty");
        if (!names.insert(c.name).second) throw std::invalid_argument("metadata column '" + c.name + "' appears twice");
        const MetaColumn* have = nullptr;
        for (const MetaColumn& m : md.columns)
            if (m.name == c.name) have = &m;
        if (c.size() != n)
            throw std::invalid_argument("metadata column '" + c.name + "' has " + std::to_string(c.size()) +
                                        " values for " + std::to_string(n) + " vectors");
        if (!c.valid.empty() && c.valid.size() != n)
            throw std::invalid_argument("metadata column '" + c.name + "' has " + std::to_string(c.valid.size()) +
                                        " null flags for " + std::to_string(n) + " vectors");
        const bool only_nulls = n > 0 && !c.valid.empty() && std::find(c.valid.begin(), c.valid.end(), 1) == c.valid.end();
        if (have && n > 0 && !only_nulls && !fits(have->type, c.type))
            throw std::invalid_argument("metadata column '" + c.name + "' holds " + kind_name(have->type) +
                                        " values; got " + kind_name(c.type));
        if (!have && only_nulls)
            throw std::invalid_argument("metadata column '" + c.name + "' is new and has only nulls, so it has no type");
    }
}

void append_meta(Metadata& md, size_t rows, size_t n, const std::vector<MetaInput>& in) {
    if (n == 0) return; // an add of no rows creates no columns: it says nothing about their types
    for (const MetaInput& c : in) {
        bool have = false;
        for (const MetaColumn& m : md.columns) have = have || m.name == c.name;
        if (have) continue;
        MetaColumn col;
        col.name = c.name;
        col.type = c.type;
        col.valid.assign(rows + n, 0);
        if (c.type == 0) col.ints.assign(rows + n, 0);
        else if (c.type == 1) col.floats.assign(rows + n, 0.0);
        else col.codes.assign(rows + n, 0);
        md.columns.push_back(std::move(col));
    }
    for (const MetaInput& c : in) {
        MetaColumn* col = nullptr;
        for (MetaColumn& m : md.columns)
            if (m.name == c.name) col = &m;
        if (col->valid.size() < rows + n) throw std::logic_error("metadata rows were not appended");
        std::unordered_map<std::string, uint32_t> code_of;
        if (col->type == 2)
            for (size_t k = 0; k < col->dict.size(); ++k) code_of.emplace(col->dict[k], static_cast<uint32_t>(k));
        for (size_t i = 0; i < n; ++i) {
            const bool present = c.valid.empty() || c.valid[i];
            const size_t r = rows + i;
            col->valid[r] = present ? 1 : 0;
            if (!present) continue;
            if (col->type == 0) {
                col->ints[r] = c.ints[i];
            } else if (col->type == 1) {
                col->floats[r] = c.type == 0 ? static_cast<double>(c.ints[i]) : c.floats[i];
            } else {
                // each new string takes the next code, in row order
                const auto ins = code_of.emplace(c.strings[i], static_cast<uint32_t>(col->dict.size()));
                if (ins.second) col->dict.push_back(c.strings[i]);
                col->codes[r] = ins.first->second;
            }
        }
    }
}

std::vector<std::pair<std::string, std::string>> meta_schema(const Metadata& md) {
    std::vector<std::pair<std::string, std::string>> out;
    for (const MetaColumn& c : md.columns) out.emplace_back(c.name, c.type == 0 ? "int" : c.type == 1 ? "float" : "str");
    return out;
}

} // namespace qh
