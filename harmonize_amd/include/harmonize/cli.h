// Command-line flags for host mains (reference util/cli.h:314-497):
// `cli::ArgSet args(argc, argv); unsigned n = args["limit"] | 1024u;`
// A flag is `-name value`; a flag given without a value reads as true when
// converted to bool; `| default` supplies the value of an absent flag.
// Not on the execution path.
#pragma once

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace util {
namespace cli {

struct ArgSet {
    int argc;
    char const** argv;

    ArgSet(int c, char const** v) : argc(c), argv(v) {}
    ArgSet(int c, char** v) : argc(c), argv((char const**)v) {}

    int get_flag_idx(char const* flag) const {
        for (int i = 0; i < argc; i++) {
            char const* s = argv[i];
            if (s != nullptr && s[0] == '-' && std::strcmp(s + 1, flag) == 0) return i;
        }
        return -1;
    }

    // The text after the flag, "" for a trailing flag, nullptr if absent.
    char const* get_flag_str(char const* flag) const {
        int idx = get_flag_idx(flag);
        if (idx < 0) return nullptr;
        if (idx == argc - 1) return "";
        return argv[idx + 1];
    }

    template <typename T>
    struct ArgVal {
        T value;
        ArgVal(T v) : value(v) {}
        ArgVal<T> operator|(T) { return *this; }
        operator T() const { return value; }
    };

    struct ArgQuery {
        char const* flag_str;
        char const* value_str;

        ArgQuery(char const* f, char const* v) : flag_str(f), value_str(v) {}

        template <typename T>
        static bool parse(char const* str, T& dest) {
            if constexpr (std::is_same<T, bool>::value) {
                if (std::strcmp(str, "true") == 0 || std::strcmp(str, "1") == 0 || str[0] == '\0') {
                    dest = true;
                    return true;
                }
                if (std::strcmp(str, "false") == 0 || std::strcmp(str, "0") == 0) {
                    dest = false;
                    return true;
                }
                return false;
            } else if constexpr (std::is_floating_point<T>::value) {
                char* end = nullptr;
                errno = 0;
                double v = std::strtod(str, &end);
                if (end == str || errno != 0) return false;
                dest = (T)v;
                return true;
            } else if constexpr (std::is_integral<T>::value && std::is_signed<T>::value) {
                char* end = nullptr;
                errno = 0;
                long long v = std::strtoll(str, &end, 0);
                if (end == str || errno != 0) return false;
                dest = (T)v;
                return true;
            } else if constexpr (std::is_integral<T>::value) {
                char* end = nullptr;
                errno = 0;
                unsigned long long v = std::strtoull(str, &end, 0);
                if (end == str || errno != 0) return false;
                dest = (T)v;
                return true;
            } else {
                return false;
            }
        }

        template <typename T>
        void scan_or_fail(T& dest) const {
            if constexpr (std::is_same<T, bool>::value) {
                if (value_str == nullptr) {
                    dest = false;
                    return;
                }
                if (!parse(value_str, dest)) dest = true;  // a bare flag followed by another flag
                return;
            } else {
                if (value_str == nullptr) {
                    std::fprintf(stderr, "No value provided for flag '-%s'\n", flag_str);
                    std::exit(1);
                }
                if (!parse(value_str, dest)) {
                    std::fprintf(stderr, "Value '%s' for flag '-%s' could not be parsed\n", value_str, flag_str);
                    std::exit(1);
                }
            }
        }

        template <typename T>
        ArgVal<T> operator|(T other) {
            if (value_str == nullptr) return ArgVal<T>(other);
            T value;
            scan_or_fail(value);
            return ArgVal<T>(value);
        }

        ArgQuery operator|(ArgQuery other) { return value_str == nullptr ? other : *this; }

        template <typename T>
        operator T() const {
            T value;
            scan_or_fail(value);
            return value;
        }
    };

    ArgQuery operator[](char const* flag_str) const { return ArgQuery(flag_str, get_flag_str(flag_str)); }
};

}  // namespace cli
}  // namespace util
