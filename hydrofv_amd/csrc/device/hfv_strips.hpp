// hfv_strips.hpp -- strip heights of the 2D step kernels (host code).
//
// A 2D step launches one wavefront per strip: 60 interior columns (STRIP_COLS_2D)
// and `rows` rows.  The device holds a fixed number of step wavefronts at once
// (its CUs times the kernel's resident workgroups per CU: 256 x 8 = 2048 on an
// MI355X at two wavefronts per SIMD), so a launch of T strips runs as about
// T / slots rounds of resident wavefronts, and a partly filled last round
// leaves SIMDs idle while it drains.
#pragma once

namespace hfv {

constexpr int STRIP_COLS_2D = 60;

// 32 rows (each strip re-derives 4 halo rows, about 12% extra work at 32),
// halved down to 8 while the grid has fewer than 4096 strips, so that a small
// grid still puts several wavefronts on each of the chip's 1024 SIMDs.
inline int default_strip_rows(int nx, int ny)
{
    const long long nsx = (nx + STRIP_COLS_2D - 1) / STRIP_COLS_2D;
    int rows = 32;
    while (rows > 8 && nsx * ((ny + rows - 1) / rows) < 4096) rows /= 2;
    return rows;
}

// Fraction of the resident slots the rounds of a launch of `strips` keep
// busy: (strips / slots) / ceil(strips / slots).
inline double strip_round_fill(long long strips, long long slots)
{
    const long long rounds = (strips + slots - 1) / slots;
    return (double)strips / ((double)rounds * (double)slots);
}

// The strip height when the caller leaves it to the library: the height near
// default_strip_rows whose launch fills its rounds of resident wavefronts
// best (ties: the height nearest the default).  Where the default is a full
// 32 rows the heights tried are [28, 40], e.g. 35 rows for 4096^2 on 2048
// slots (8142 strips, 99.4% of four rounds, where 32 rows give 8832 strips,
// 86% of five); below that, the default +-25%, e.g. 18 rows for 2048^2
// (3990 strips, 97% of two rounds, where 16 rows give 4480, 73% of three).
// slots <= 0 (unknown): the default.
inline int balanced_strip_rows(int nx, int ny, long long slots)
{
    const int base = default_strip_rows(nx, ny);
    if (slots <= 0) return base;
    const int lo = (base == 32) ? 28 : base - base / 4, hi = (base == 32) ? 40 : base + base / 4;
    const long long nsx = (nx + STRIP_COLS_2D - 1) / STRIP_COLS_2D;
    int best = base;
    double best_fill = strip_round_fill(nsx * ((ny + base - 1) / base), slots);
    for (int r = lo; r <= hi; r++) {
        const double f = strip_round_fill(nsx * ((ny + r - 1) / r), slots);
        const int dr = r > base ? r - base : base - r, db = best > base ? best - base : base - best;
        if (f > best_fill + 1e-3 || (f >= best_fill - 1e-3 && dr < db)) {
            best = r;
            best_fill = f;
        }
    }
    return best;
}

// The exact solver's default (its split form, mh2d.hip): most of its strip
// pass is tiles of one state, whose cost is their loads and stores, so fewer,
// taller strips (fewer halo rows re-read, fewer rounds) pay where a grid has
// at least half a round of 32-row strips: the height in [28, 40] that fills
// its rounds best (ties: nearest 32), e.g. 36 rows for 2048^2 on 2048 slots
// (1995 strips, 97% of one round; the Sedov run 2.68 -> 2.60 s, where 18 rows
// are the general default).  Smaller grids take balanced_strip_rows.
inline int exact_strip_rows(int nx, int ny, long long slots)
{
    const long long nsx = (nx + STRIP_COLS_2D - 1) / STRIP_COLS_2D;
    if (slots <= 0 || 2 * nsx * ((ny + 31) / 32) < slots) return balanced_strip_rows(nx, ny, slots);
    int best = 32;
    double best_fill = strip_round_fill(nsx * ((ny + 31) / 32), slots);
    for (int r = 28; r <= 40; r++) {
        const double f = strip_round_fill(nsx * ((ny + r - 1) / r), slots);
        const int dr = r > 32 ? r - 32 : 32 - r, db = best > 32 ? best - 32 : 32 - best;
        if (f > best_fill + 1e-3 || (f >= best_fill - 1e-3 && dr < db)) {
            best = r;
            best_fill = f;
        }
    }
    return best;
}

}  // namespace hfv
