// cli_args.hpp -- command-line flags with the semantics the reference CLIs
// rely on (src/grid/grid.cpp, src/dft/dft.cpp):
//   - any number of leading '-' is ignored and names compare case-insensitively;
//   - a presence flag matches when the name part (up to '=') has exactly the
//     flag's length;
//   - a value lookup matches any argument that starts with the name, takes the
//     text after the name (skipping one '=' if present), last match wins, and
//     yields 0 when the flag is absent.
#pragma once

#include <cstdlib>
#include <cstring>
#include <strings.h>

namespace cli {

inline const char *strip_dashes(const char *a)
{
    const size_t len = strlen(a);
    size_t k = 0;
    while (a[k] == '-')
        ++k;
    // a run of dashes covering (almost) the whole argument is kept as is
    if (len == 0 || k + 1 >= len)
        return a;
    return a + k;
}

inline bool has_flag(int argc, char **argv, const char *name)
{
    const size_t n = strlen(name);
    bool found = false;
    for (int i = 1; i < argc; ++i) {
        const char *s = strip_dashes(argv[i]);
        const char *eq = strchr(s, '=');
        const size_t len = eq ? (size_t)(eq - s) : strlen(s);
        if (len == n && strncasecmp(s, name, n) == 0)
            found = true;
    }
    return found;
}

// text of the last argument starting with `name`, nullptr when none
inline const char *value_text(int argc, char **argv, const char *name, bool skip_any)
{
    const size_t n = strlen(name);
    const char *val = nullptr;
    for (int i = 1; i < argc; ++i) {
        const char *s = strip_dashes(argv[i]);
        if (strncasecmp(s, name, n) != 0)
            continue;
        if (skip_any)
            val = strlen(s) > n ? s + n + 1 : s + n;
        else
            val = s + n + (s[n] == '=' ? 1 : 0);
    }
    return val;
}

inline double get_double(int argc, char **argv, const char *name)
{
    const char *t = value_text(argc, argv, name, false);
    return t ? atof(t) : 0.0;
}

inline float get_float(int argc, char **argv, const char *name)
{
    return (float)get_double(argc, argv, name);
}

inline int get_int(int argc, char **argv, const char *name)
{
    const char *t = value_text(argc, argv, name, false);
    return t ? atoi(t) : 0;
}

inline const char *get_string(int argc, char **argv, const char *name)
{
    return value_text(argc, argv, name, true);
}

} // namespace cli
