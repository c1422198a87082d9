// dcm_namelist.cpp -- a reader of Fortran namelist files, for the model's
// run and sounding namelists (main.f90:74-80 `read(8,nml=run_vars)` and
// :134-140 `read(8,nml=sounding_vars)`; namelist.in).
//
// Supported: groups `&name ... /`, case-insensitive object names with
// derived-type components (`nm1%dt`), subscripts with single indices and
// ranges (`q_read(1,1:7)`), value lists with repeat counts (`3*0.`) and
// null values (consecutive commas leave an element unchanged), reals in
// Fortran notation (`1.e-8`, `2.5d0`), integers, logicals (`.true.`, `T`,
// `.f.`), character values in apostrophes or quotes (doubled delimiters
// inside), and `!` comments.  Values are kept as text and converted on
// request, as the Fortran read converts to the type of the object.
#include <cctype>
#include <cerrno>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dcm_hip.h"

namespace dcm {
void set_error(const std::string &msg);
}

namespace {

struct Value {
    bool null = false;  // an empty list element
    bool quoted = false;
    std::string text;
};

struct Assign {
    std::string name;                       // lower case, with components
    std::vector<std::pair<long, long>> sub; // (lo, hi) per subscript; empty: none
    std::vector<Value> values;              // repeat counts expanded
};

struct Group {
    std::string name;
    std::vector<Assign> items;
};

std::string lower(std::string s) {
    for (char &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

class Parser {
public:
    explicit Parser(const std::string &t) : s(t) {}
    bool parse(std::vector<Group> &out, std::string *why);

private:
    const std::string &s;
    size_t p = 0;
    std::string err;

    void skip_ws() {
        for (;;) {
            while (p < s.size() && (std::isspace((unsigned char)s[p]))) ++p;
            if (p < s.size() && s[p] == '!') {
                while (p < s.size() && s[p] != '\n') ++p;
                continue;
            }
            return;
        }
    }
    bool is_name_char(char c) { return std::isalnum((unsigned char)c) || c == '_' || c == '%'; }
    std::string read_name() {
        size_t b = p;
        while (p < s.size() && is_name_char(s[p])) ++p;
        return lower(s.substr(b, p - b));
    }
    bool read_subscripts(std::vector<std::pair<long, long>> &sub);
    bool read_values(std::vector<Value> &vals);
    // true if the text at p starts a new `name =` or `name(...) =`
    bool at_assignment();
};

bool Parser::read_subscripts(std::vector<std::pair<long, long>> &sub) {
    // at '('
    ++p;
    for (;;) {
        skip_ws();
        char *end = nullptr;
        const long lo = std::strtol(s.c_str() + p, &end, 10);
        if (end == s.c_str() + p) return false;
        p = end - s.c_str();
        long hi = lo;
        skip_ws();
        if (p < s.size() && s[p] == ':') {
            ++p;
            skip_ws();
            hi = std::strtol(s.c_str() + p, &end, 10);
            if (end == s.c_str() + p) return false;
            p = end - s.c_str();
            skip_ws();
        }
        sub.push_back({lo, hi});
        if (p < s.size() && s[p] == ',') {
            ++p;
            continue;
        }
        if (p < s.size() && s[p] == ')') {
            ++p;
            return true;
        }
        return false;
    }
}

bool Parser::at_assignment() {
    size_t q = p;
    if (q >= s.size() || !(std::isalpha((unsigned char)s[q]))) return false;
    while (q < s.size() && is_name_char(s[q])) ++q;
    while (q < s.size() && std::isspace((unsigned char)s[q])) ++q;
    if (q < s.size() && s[q] == '(') {
        int depth = 0;
        for (; q < s.size(); ++q) {
            if (s[q] == '(') ++depth;
            if (s[q] == ')' && --depth == 0) {
                ++q;
                break;
            }
        }
        while (q < s.size() && std::isspace((unsigned char)s[q])) ++q;
    }
    return q < s.size() && s[q] == '=';
}

bool Parser::read_values(std::vector<Value> &vals) {
    bool expect_value = true;  // a separator without a value in between is a null
    for (;;) {
        skip_ws();
        if (p >= s.size()) return true;
        const char c = s[p];
        if (c == '/' || c == '&' || c == '$') return true;
        if (c == ',') {
            if (expect_value) {
                Value v;
                v.null = true;
                vals.push_back(v);
            }
            expect_value = true;
            ++p;
            continue;
        }
        if (at_assignment()) return true;
        // one value, possibly r*value
        long repeat = 1;
        {
            size_t q = p;
            while (q < s.size() && std::isdigit((unsigned char)s[q])) ++q;
            if (q > p && q < s.size() && s[q] == '*') {
                repeat = std::strtol(s.substr(p, q - p).c_str(), nullptr, 10);
                p = q + 1;
            }
        }
        Value v;
        if (p < s.size() && (s[p] == '\'' || s[p] == '"')) {
            const char d = s[p++];
            v.quoted = true;
            for (;;) {
                if (p >= s.size()) {
                    err = "unterminated character value";
                    return false;
                }
                if (s[p] == d) {
                    if (p + 1 < s.size() && s[p + 1] == d) {
                        v.text += d;
                        p += 2;
                        continue;
                    }
                    ++p;
                    break;
                }
                v.text += s[p++];
            }
        } else {
            size_t b = p;
            while (p < s.size() && !std::isspace((unsigned char)s[p]) && s[p] != ',' &&
                   s[p] != '/' && s[p] != '!')
                ++p;
            v.text = s.substr(b, p - b);
            if (v.text.empty()) v.null = true;
        }
        for (long r = 0; r < repeat; ++r) vals.push_back(v);
        expect_value = false;
    }
}

bool Parser::parse(std::vector<Group> &out, std::string *why) {
    for (;;) {
        skip_ws();
        if (p >= s.size()) return true;
        if (s[p] != '&' && s[p] != '$') {
            // text outside a group is ignored, as the Fortran read skips it
            while (p < s.size() && s[p] != '\n') ++p;
            continue;
        }
        ++p;
        Group g;
        g.name = read_name();
        if (g.name == "end") continue;
        for (;;) {
            skip_ws();
            if (p >= s.size()) {
                if (why) *why = "group &" + g.name + " has no terminating /";
                return false;
            }
            if (s[p] == '/') {
                ++p;
                break;
            }
            if (s[p] == '&' || s[p] == '$') {  // &end / $end
                ++p;
                read_name();
                break;
            }
            Assign a;
            a.name = read_name();
            if (a.name.empty()) {
                if (why) *why = "bad object name in group &" + g.name + " near: " +
                                s.substr(p, 30);
                return false;
            }
            skip_ws();
            if (p < s.size() && s[p] == '(' && !read_subscripts(a.sub)) {
                if (why) *why = "bad subscripts for " + a.name;
                return false;
            }
            skip_ws();
            if (p >= s.size() || s[p] != '=') {
                if (why) *why = "expected = after " + a.name;
                return false;
            }
            ++p;
            if (!read_values(a.values)) {
                if (why) *why = err + " in " + a.name;
                return false;
            }
            g.items.push_back(a);
        }
        out.push_back(g);
    }
}

bool to_real(const Value &v, double *out) {
    if (v.null || v.quoted) return false;
    std::string t = v.text;
    for (char &c : t)
        if (c == 'd' || c == 'D' || c == 'q' || c == 'Q') c = 'e';
    char *end = nullptr;
    errno = 0;
    const double x = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0' || errno == ERANGE) return false;
    *out = x;
    return true;
}

bool to_logical(const Value &v, int *out) {
    if (v.null || v.quoted) return false;
    std::string t = lower(v.text);
    if (!t.empty() && t[0] == '.') t = t.substr(1);
    if (t.empty()) return false;
    if (t[0] == 't') {
        *out = 1;
        return true;
    }
    if (t[0] == 'f') {
        *out = 0;
        return true;
    }
    return false;
}

}  // namespace

struct dcm_namelist {
    std::vector<Group> groups;
    const Group *group(const char *name) const {
        const std::string n = lower(name ? name : "");
        for (const Group &g : groups)
            if (g.name == n) return &g;
        return nullptr;
    }
    // the last assignment of a scalar (no subscripts, or a single element)
    const Value *scalar(const char *grp, const char *name) const {
        const Group *g = group(grp);
        if (!g || !name) return nullptr;
        const std::string n = lower(name);
        const Value *found = nullptr;
        for (const Assign &a : g->items)
            if (a.name == n && !a.values.empty() && !a.values[0].null) found = &a.values[0];
        return found;
    }
};

extern "C" {

int dcm_nml_read(const char *path, dcm_namelist **out) {
    if (!path || !out) return DCM_BAD_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f) {
        dcm::set_error(std::string("dcm_nml_read: cannot open ") + path);
        return DCM_IO_ERROR;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    dcm_namelist *nl = new dcm_namelist;
    std::string why;
    if (!Parser(text).parse(nl->groups, &why)) {
        delete nl;
        dcm::set_error(std::string("dcm_nml_read: ") + path + ": " + why);
        return DCM_BAD_ARG;
    }
    *out = nl;
    return DCM_OK;
}

int dcm_nml_free(dcm_namelist *nl) {
    delete nl;
    return DCM_OK;
}

int dcm_nml_has(const dcm_namelist *nl, const char *group, const char *name) {
    if (!nl) return 0;
    const Group *g = nl->group(group);
    if (!g) return 0;
    if (!name) return 1;
    const std::string n = lower(name);
    for (const Assign &a : g->items)
        if (a.name == n) return 1;
    return 0;
}

int dcm_nml_get_real(const dcm_namelist *nl, const char *group, const char *name, double *v) {
    if (!nl || !v) return DCM_BAD_ARG;
    const Value *x = nl->scalar(group, name);
    if (!x || !to_real(*x, v)) {
        dcm::set_error(std::string("dcm_nml_get_real: no real ") + (name ? name : ""));
        return DCM_BAD_ARG;
    }
    return DCM_OK;
}

int dcm_nml_get_int(const dcm_namelist *nl, const char *group, const char *name, int *v) {
    if (!nl || !v) return DCM_BAD_ARG;
    const Value *x = nl->scalar(group, name);
    if (x && !x->null && !x->quoted) {
        char *end = nullptr;
        const long r = std::strtol(x->text.c_str(), &end, 10);
        if (end != x->text.c_str() && *end == '\0') {
            *v = (int)r;
            return DCM_OK;
        }
    }
    dcm::set_error(std::string("dcm_nml_get_int: no integer ") + (name ? name : ""));
    return DCM_BAD_ARG;
}

int dcm_nml_get_logical(const dcm_namelist *nl, const char *group, const char *name, int *v) {
    if (!nl || !v) return DCM_BAD_ARG;
    const Value *x = nl->scalar(group, name);
    if (!x || !to_logical(*x, v)) {
        dcm::set_error(std::string("dcm_nml_get_logical: no logical ") + (name ? name : ""));
        return DCM_BAD_ARG;
    }
    return DCM_OK;
}

int dcm_nml_get_string(const dcm_namelist *nl, const char *group, const char *name, char *buf,
                       size_t n) {
    if (!nl || !buf || n == 0) return DCM_BAD_ARG;
    const Value *x = nl->scalar(group, name);
    if (!x) {
        dcm::set_error(std::string("dcm_nml_get_string: no value ") + (name ? name : ""));
        return DCM_BAD_ARG;
    }
    std::strncpy(buf, x->text.c_str(), n - 1);
    buf[n - 1] = '\0';
    return DCM_OK;
}

int dcm_nml_get_array(const dcm_namelist *nl, const char *group, const char *name, int rank,
                      const int *shape, double *out) {
    if (!nl || !name || rank < 1 || rank > 7 || !shape || !out) return DCM_BAD_ARG;
    const Group *g = nl->group(group);
    if (!g) {
        dcm::set_error(std::string("dcm_nml_get_array: no group ") + (group ? group : ""));
        return DCM_BAD_ARG;
    }
    const std::string n = lower(name);
    long total = 1;
    for (int d = 0; d < rank; ++d) total *= shape[d];
    bool any = false;
    for (const Assign &a : g->items) {
        if (a.name != n) continue;
        any = true;
        // the section in array-element (column-major) order; a missing or
        // short subscript list starts at that element and runs on
        std::vector<long> lo(rank, 1), hi(rank, 1);
        long start = 0;  // linear start when the list fills on from an element
        bool section = (int)a.sub.size() == rank;
        if (a.sub.empty()) {
            for (int d = 0; d < rank; ++d) hi[d] = shape[d];
            section = true;
        }
        std::vector<long> idx;
        if (section) {
            for (int d = 0; d < rank && !a.sub.empty(); ++d) {
                lo[d] = a.sub[d].first;
                hi[d] = a.sub[d].second;
            }
            for (int d = 0; d < rank; ++d)
                if (lo[d] < 1 || hi[d] > shape[d] || lo[d] > hi[d]) {
                    dcm::set_error("dcm_nml_get_array: subscript out of bounds in " + n);
                    return DCM_BAD_ARG;
                }
            const bool single = [&] {
                for (int d = 0; d < rank; ++d)
                    if (lo[d] != hi[d]) return false;
                return true;
            }();
            if (single && a.values.size() > 1) {
                // a(i) = v1, v2, ... fills on from element i
                long e = 0, stride = 1;
                for (int d = 0; d < rank; ++d) {
                    e += (lo[d] - 1) * stride;
                    stride *= shape[d];
                }
                start = e;
                for (long t = 0; t < (long)a.values.size(); ++t) idx.push_back(start + t);
            } else {
                std::vector<long> c(lo);
                for (;;) {
                    long e = 0, stride = 1;
                    for (int d = 0; d < rank; ++d) {
                        e += (c[d] - 1) * stride;
                        stride *= shape[d];
                    }
                    idx.push_back(e);
                    int d = 0;
                    while (d < rank && c[d] == hi[d]) c[d] = lo[d], ++d;
                    if (d == rank) break;
                    ++c[d];
                }
            }
        } else {
            dcm::set_error("dcm_nml_get_array: subscripts of " + n + " do not match its rank");
            return DCM_BAD_ARG;
        }
        if (a.values.size() > idx.size()) {
            dcm::set_error("dcm_nml_get_array: too many values for " + n);
            return DCM_BAD_ARG;
        }
        for (size_t t = 0; t < a.values.size(); ++t) {
            if (a.values[t].null) continue;
            if (idx[t] >= total) {
                dcm::set_error("dcm_nml_get_array: values run past the end of " + n);
                return DCM_BAD_ARG;
            }
            double x;
            if (!to_real(a.values[t], &x)) {
                dcm::set_error("dcm_nml_get_array: a value of " + n + " is not a number");
                return DCM_BAD_ARG;
            }
            out[idx[t]] = x;
        }
    }
    if (!any) {
        dcm::set_error("dcm_nml_get_array: no " + n);
        return DCM_BAD_ARG;
    }
    return DCM_OK;
}

}  // extern "C"
