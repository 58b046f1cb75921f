#include "dict.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

namespace mplb {

bool Dict::open(const std::string& path) {
    std::ifstream in(path);
    if (!in) return false;
    std::ostringstream os;
    std::string word;
    // openDict.c reads words with fscanf("%s") into a 100-byte buffer
    while (in >> word) os << word << ' ';
    content_ = os.str();
    return true;
}

namespace {

// bracedEntry.c: first occurrence of name followed by optional spaces and '{'; the
// block content drops the "{ " and " }" around it.
bool braced(const std::string& name, const std::string& content, std::string* out) {
    size_t pos = content.find(name);
    while (pos != std::string::npos) {
        size_t p = pos + name.size();
        while (p < content.size() && (content[p] == ' ' || content[p] == '\n')) p++;
        if (p < content.size() && content[p] == '{') {
            int depth = 1;
            size_t j = 1;
            const size_t len = content.size() - p;
            while (j < len && depth != 0) {
                const char c = content[p + j];
                if (c == '{') depth++;
                else if (c == '}') depth--;
                j++;
            }
            if (depth != 0) return false;
            // copy j-4 characters starting two past the brace
            const size_t n = j >= 4 ? j - 4 : 0;
            *out = content.substr(p + 2, n);
            return true;
        }
        pos = content.find(name, pos + name.size());
    }
    return false;
}

// singleEntry.c: value = text from one past the name up to the next ';'
bool single(const std::string& name, const std::string& content, std::string* out) {
    const size_t pos = content.find(name);
    if (pos == std::string::npos) return false;
    const size_t start = pos + name.size() + 1;
    if (start > content.size()) return false;
    const size_t semi = content.find(';', start);
    if (semi == std::string::npos) return false;
    *out = content.substr(start, semi - start);
    return true;
}

}  // namespace

bool Dict::lookup(const std::string& ename, std::string* out) const {
    std::vector<std::string> tk;
    size_t s = 0;
    for (;;) {
        const size_t q = ename.find('/', s);
        if (q == std::string::npos) {
            tk.push_back(ename.substr(s));
            break;
        }
        tk.push_back(ename.substr(s, q - s));
        s = q + 1;
    }
    std::string cur = content_;
    for (size_t i = 0; i + 1 < tk.size(); i++) {
        std::string inner;
        if (!braced(tk[i], cur, &inner)) return false;
        cur = inner;
    }
    return single(tk.back(), cur, out);
}

bool Dict::scalar(const std::string& ename, double df, double* out) const {
    std::string v;
    if (!lookup(ename, &v)) {
        *out = df;
        return false;
    }
    *out = std::strtod(v.c_str(), nullptr);
    return true;
}

bool Dict::integer(const std::string& ename, int df, int* out) const {
    std::string v;
    if (!lookup(ename, &v)) {
        *out = df;
        return false;
    }
    *out = std::atoi(v.c_str());
    return true;
}

bool Dict::string_list(const std::string& ename, std::vector<std::string>* out) const {
    out->clear();
    std::string cur = content_;
    size_t s = 0;
    for (;;) {
        const size_t e = ename.find('/', s);
        const std::string tk = ename.substr(s, e == std::string::npos ? std::string::npos : e - s);
        std::string inner;
        if (!braced(tk, cur, &inner)) return false;
        cur = inner;
        if (e == std::string::npos) break;
        s = e + 1;
    }
    size_t i = 0;
    while (i < cur.size()) {
        while (i < cur.size() && (cur[i] == ' ' || cur[i] == '\n')) i++;
        size_t j = i;
        while (j < cur.size() && cur[j] != ' ' && cur[j] != '\n') j++;
        if (j > i) out->push_back(cur.substr(i, j - i));
        i = j;
    }
    return true;
}

bool Dict::vector(const std::string& ename, std::vector<double>* out) const {
    std::string v;
    out->clear();
    if (!lookup(ename, &v)) return false;
    for (char& c : v)
        if (c == '(' || c == ')') c = ' ';
    const char* p = v.c_str();
    char* end = nullptr;
    for (;;) {
        while (*p == ' ') p++;
        if (!*p) break;
        const double x = std::strtod(p, &end);
        if (end == p) {  // a non-numeric token reads as 0 (strtod on the token)
            out->push_back(0.0);
            while (*p && *p != ' ') p++;
            continue;
        }
        out->push_back(x);
        p = end;
    }
    return true;
}

bool Dict::string(const std::string& ename, const std::string& df, std::string* out) const {
    std::string v;
    if (!lookup(ename, &v)) {
        *out = df;
        return false;
    }
    *out = v;
    return true;
}

}  // namespace mplb
