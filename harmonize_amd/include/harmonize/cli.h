// Command-line utilities for host mains (reference util/cli.h):
//  - flags (util/cli.h:314-497):
//    `cli::ArgSet args(argc, argv); unsigned n = args["limit"] | 1024u;`
//    A flag is `-name value`; a flag given without a value reads as true
//    when converted to bool; `| default` supplies the value of an absent flag.
//  - text plots (util/cli.h:13-71, :93): `cli_graph(data, size, shape, style)`
//    draws a series as a bar plot on the terminal with block-element tiles.
// Not on the execution path.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
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

// A tile set for cli_graph. One character cell shows `cols` sub-columns side
// by side, each filled from the bottom to a level in [0, rows]; the cell for
// levels (l_0, ..., l_{cols-1}) is tiles[sum_k l_k * (rows+1)^(cols-1-k)].
struct GraphStyle {
    int rows;
    int cols;
    const char** tiles;
    GraphStyle(unsigned int r, unsigned int c, const char** t) : rows((int)r), cols((int)c), tiles(t) {}
};

// Plot size in character cells and the value window it shows.
struct GraphShape {
    int width;
    int height;
    float x_min;
    float x_max;
    float y_min;
    float y_max;
};

// Eighth blocks: one sub-column, eight levels per cell.
static const char* EighthBlockTiles[] = {" ", "\u2581", "\u2582", "\u2583", "\u2584",
                                         "\u2585", "\u2586", "\u2587", "\u2588"};
// Quadrant blocks: two sub-columns (left, right) of two levels each.
static const char* QuadrantTiles[] = {
    " ",      "\u2597", "\u2590",  // left empty:  right 0, 1, 2
    "\u2596", "\u2584", "\u259F",  // left lower half
    "\u258C", "\u2599", "\u2588",  // left full
};
static const GraphStyle BlockFill = GraphStyle(8, 1, EighthBlockTiles);
static const GraphStyle Block2x2Fill = GraphStyle(2, 2, QuadrantTiles);

// Draws data[0, size) as `shape.height` text rows of `shape.width` cells.
// Each sub-column shows the mean of the samples that fall into it (an empty
// one repeats its left neighbour); the y axis is labelled per row and the x
// axis every fourth cell, both from the shape's window.
inline void cli_graph(float* data, int size, GraphShape shape, GraphStyle style) {
    if (data == nullptr || size <= 0 || shape.width <= 0 || shape.height <= 0 || style.rows <= 0 ||
        style.cols <= 0) {
        return;
    }
    const int subs = shape.width * style.cols;
    float* means = new float[subs];
    float carry = data[0];
    for (int s = 0; s < subs; ++s) {
        long long lo = (long long)s * size / subs;
        long long hi = (long long)(s + 1) * size / subs;
        if (hi > lo) {
            double acc = 0.0;
            for (long long i = lo; i < hi; ++i) acc += data[i];
            carry = (float)(acc / (double)(hi - lo));
        }
        means[s] = carry;
    }
    const float span = shape.y_max - shape.y_min;
    const float band = span / (float)shape.height;
    for (int row = 0; row < shape.height; ++row) {
        const float floor_val = shape.y_min + band * (float)(shape.height - 1 - row);
        std::printf("%11.4g |", (double)floor_val);
        for (int c = 0; c < shape.width; ++c) {
            int index = 0;
            for (int k = 0; k < style.cols; ++k) {
                float v = means[c * style.cols + k];
                int level = 0;
                if (band > 0.0f) {
                    float rel = (v - floor_val) / band * (float)style.rows;
                    level = rel <= 0.0f ? 0 : (rel >= (float)style.rows ? style.rows : (int)rel);
                } else if (v >= floor_val) {
                    level = style.rows;
                }
                index = index * (style.rows + 1) + level;
            }
            std::fputs(style.tiles[index], stdout);
        }
        std::printf("\n");
    }
    std::printf("%11s +", "");
    for (int c = 0; c < shape.width; ++c) std::fputs(c % 4 == 0 ? "+" : "-", stdout);
    std::printf("\n%13s", "");
    std::string axis;
    for (int c = 0; c < shape.width; c += 4) {
        char label[32];
        float x = shape.x_min + (shape.x_max - shape.x_min) * (float)c / (float)shape.width;
        std::snprintf(label, sizeof(label), "%-4.3g", (double)x);
        std::string cell(label);
        cell.resize(4, ' ');
        axis += cell.substr(0, 4);
    }
    std::printf("%s\n", axis.c_str());
    delete[] means;
}

}  // namespace cli
}  // namespace util
