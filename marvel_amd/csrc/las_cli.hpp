// Helpers shared by the .las utility commands: the reference's PathTo / Root
// (db/DB.c:87-140) and Print_Number (db/DB.c: comma-grouped integers).
#pragma once
#include <strings.h>

#include <cstdio>
#include <cstring>
#include <string>

namespace lascli {

// directory part of name, "." without one
inline std::string path_to(const std::string& name)
{
    const size_t s = name.rfind('/');
    return s == std::string::npos ? "." : name.substr(0, s);
}

// file part of name with a trailing `suffix` (any case) removed when it leaves something
inline std::string root_of(const std::string& name, const char* suffix)
{
    const size_t s = name.rfind('/');
    std::string f = s == std::string::npos ? name : name.substr(s + 1);
    const size_t n = strlen(suffix);
    if (f.size() > n && strcasecmp(f.c_str() + f.size() - n, suffix) == 0) f.resize(f.size() - n);
    return f;
}

inline void print_number(long long v, FILE* out)
{
    std::string d = std::to_string(v < 0 ? -v : v), r;
    for (size_t i = 0; i < d.size(); i++) {
        if (i && (d.size() - i) % 3 == 0) r += ',';
        r += d[i];
    }
    fprintf(out, "%s%s", v < 0 ? "-" : "", r.c_str());
}

}  // namespace lascli
