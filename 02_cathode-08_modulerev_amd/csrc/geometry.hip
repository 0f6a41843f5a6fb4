// geometry.hip -- the plasma region of the implicit NDD step as block boxes
// and boundary segments (host side): validation, per-face type codes, ghost
// rules, and the reference's five-block cathode
// (include/initialModule.hpp:239-343 for the boxes, the boundary loops of
// fluidModule.hpp:2152-2249, 3873-3913 and 4906-4946 for the segments).
#include <algorithm>
#include <cstring>
#include <string>

#include "common.hpp"
#include "geometry.hpp"

namespace cathode {

cathode_ndd_geometry rect_geometry(int ni, int nj)
{
    cathode_ndd_geometry g;
    std::memset(&g, 0, sizeof(g));
    g.n_bl = 1;
    const int c[4] = {1, ni, 1, nj}, x[4] = {2, ni, 1, nj}, r[4] = {1, ni, 2, nj};
    for (int q = 0; q < 4; q++) {
        g.cell_box[q] = c[q];
        g.xface_box[q] = x[q];
        g.rface_box[q] = r[q];
    }
    g.n_seg = 4;
    g.seg[0] = {CATHODE_SEG_X_LOW, 1, 1, nj, 0};
    g.seg[1] = {CATHODE_SEG_X_HIGH, ni, 1, nj, 1};
    g.seg[2] = {CATHODE_SEG_R_OPEN, nj, 1, ni, 0};
    g.seg[3] = {CATHODE_SEG_R_AXIS, 1, 1, ni, 0};
    g.n_ne = 0;
    g.n_te = 4;
    g.te[0] = {CATHODE_SEG_X_LOW, 1, 1, nj, 1};
    g.te[1] = {CATHODE_SEG_X_HIGH, ni, 1, nj, 1};
    g.te[2] = {CATHODE_SEG_R_LOW, 1, 1, ni, 1};
    g.te[3] = {CATHODE_SEG_R_HIGH, nj, 1, ni, 1};
    return g;
}

namespace {

bool box_in(const int* b, int i0, int i1, int j0, int j1)
{
    return b[0] <= b[1] && b[2] <= b[3] && b[0] >= i0 && b[1] <= i1 && b[2] >= j0 && b[3] <= j1;
}

bool boxes_overlap(const int* a, const int* b)
{
    return !(a[1] < b[0] || b[1] < a[0] || a[3] < b[2] || b[3] < a[2]);
}

// the plasma cell, the face (as (i, j) of the face array) and the neighbour
// across the face of segment kind k at cell index c and running index t
struct SegPoint {
    int ci, cj, fi, fj, gi, gj;
};
SegPoint seg_point(int kind, int c, int t)
{
    switch (kind) {
    case CATHODE_SEG_X_LOW: return {c, t, c, t, c - 1, t};
    case CATHODE_SEG_X_HIGH: return {c, t, c + 1, t, c + 1, t};
    case CATHODE_SEG_R_LOW:
    case CATHODE_SEG_R_AXIS: return {t, c, t, c, t, c - 1};
    default: return {t, c, t, c + 1, t, c + 1};  // R_HIGH, R_OPEN
    }
}

int side_bit(int kind)
{
    switch (kind) {
    case CATHODE_SEG_X_LOW: return 0;
    case CATHODE_SEG_X_HIGH: return 1;
    case CATHODE_SEG_R_LOW: return 2;
    default: return 3;
    }
}

}  // namespace

bool build_geo_tables(int ni, int nj, const cathode_ndd_geometry& g, GeoTables& T, std::string& err)
{
    auto fail = [&](const std::string& m) {
        err = "ndd_set_geometry: " + m;
        return false;
    };
    if (g.n_bl < 1 || g.n_bl > CATHODE_NDD_MAX_BLOCKS) return fail("n_bl out of range");
    if (g.n_seg < 0 || g.n_seg > CATHODE_NDD_MAX_SEGS || g.n_ne < 0 || g.n_ne > CATHODE_NDD_MAX_SEGS || g.n_te < 0 ||
        g.n_te > CATHODE_NDD_MAX_SEGS)
        return fail("segment count out of range");
    const int ld = nj + 2;
    const size_t n = (size_t)(ni + 2) * ld;
    T.ni = ni;
    T.nj = nj;
    T.ld = ld;
    T.xc.assign(n, GF_NONE);
    T.rc.assign(n, GF_NONE);
    T.act.assign(n, 0);
    T.gs.assign(n, -1);
    T.ne.assign(n, 0);
    T.te.assign(n, 0);
    T.blk.assign(n, 0);
    T.anode.clear();
    auto at = [&](int i, int j) { return (size_t)i * ld + j; };
    // boxes: inside the grid and disjoint within each kind
    for (int b = 0; b < g.n_bl; b++) {
        const int *c = g.cell_box + 4 * b, *x = g.xface_box + 4 * b, *r = g.rface_box + 4 * b;
        if (!box_in(c, 1, ni, 1, nj)) return fail("cell box " + std::to_string(b) + " outside 1..ni x 1..nj");
        // an empty face box is allowed (i0 > i1 or j0 > j1 only when the block has one cell row / column)
        if (x[0] <= x[1] && x[2] <= x[3] && !box_in(x, 2, ni, 1, nj))
            return fail("x-face box " + std::to_string(b) + " outside the inner faces 2..ni x 1..nj");
        if (r[0] <= r[1] && r[2] <= r[3] && !box_in(r, 1, ni, 2, nj))
            return fail("r-face box " + std::to_string(b) + " outside the inner faces 1..ni x 2..nj");
        for (int a = 0; a < b; a++) {
            if (boxes_overlap(g.cell_box + 4 * a, c)) return fail("cell boxes overlap");
            const int *xa = g.xface_box + 4 * a, *ra = g.rface_box + 4 * a;
            if (x[0] <= x[1] && x[2] <= x[3] && xa[0] <= xa[1] && xa[2] <= xa[3] && boxes_overlap(xa, x))
                return fail("x-face boxes overlap");
            if (r[0] <= r[1] && r[2] <= r[3] && ra[0] <= ra[1] && ra[2] <= ra[3] && boxes_overlap(ra, r))
                return fail("r-face boxes overlap");
        }
        for (int i = c[0]; i <= c[1]; i++)
            for (int j = c[2]; j <= c[3]; j++) {
                T.act[at(i, j)] = 1;
                T.blk[at(i, j)] = (uint8_t)(b + 1);
            }
    }
    T.ncells = 0;
    for (size_t k = 0; k < n; k++) T.ncells += T.act[k];
    // inner faces: between two plasma cells
    for (int b = 0; b < g.n_bl; b++) {
        const int *x = g.xface_box + 4 * b, *r = g.rface_box + 4 * b;
        for (int i = x[0]; i <= x[1]; i++)
            for (int j = x[2]; j <= x[3]; j++) {
                if (!T.act[at(i, j)] || !T.act[at(i - 1, j)])
                    return fail("x face (" + std::to_string(i) + "," + std::to_string(j) +
                                ") of a face box does not lie between two plasma cells");
                T.xc[at(i, j)] = GF_INNER;
            }
        for (int i = r[0]; i <= r[1]; i++)
            for (int j = r[2]; j <= r[3]; j++) {
                if (!T.act[at(i, j)] || !T.act[at(i, j - 1)])
                    return fail("r face (" + std::to_string(i) + "," + std::to_string(j) +
                                ") of a face box does not lie between two plasma cells");
                T.rc[at(i, j)] = GF_INNER;
            }
    }
    // boundary segments, in order: face codes and the potential ghost rules
    // (a later segment's ghost write replaces an earlier one, as in the
    // reference's loops)
    for (int s = 0; s < g.n_seg; s++) {
        const cathode_ndd_segment& q = g.seg[s];
        if (q.kind < CATHODE_SEG_X_LOW || q.kind > CATHODE_SEG_R_AXIS || q.lo > q.hi)
            return fail("segment " + std::to_string(s) + ": bad kind or range");
        const bool xk = q.kind == CATHODE_SEG_X_LOW || q.kind == CATHODE_SEG_X_HIGH;
        if (xk ? (q.cell < 1 || q.cell > ni || q.lo < 1 || q.hi > nj) : (q.cell < 1 || q.cell > nj || q.lo < 1 || q.hi > ni))
            return fail("segment " + std::to_string(s) + " outside the grid");
        if ((q.flag & ~(CATHODE_SEG_BIAS | CATHODE_SEG_INLET | CATHODE_SEG_UIX_BELOW)) ||
            ((q.flag & CATHODE_SEG_BIAS) && !xk) ||
            ((q.flag & (CATHODE_SEG_INLET | CATHODE_SEG_UIX_BELOW)) && q.kind != CATHODE_SEG_R_HIGH))
            return fail("segment " + std::to_string(s) +
                        ": flag bits are CATHODE_SEG_BIAS on x walls, CATHODE_SEG_INLET and "
                        "CATHODE_SEG_UIX_BELOW on high-r walls");
        uint8_t code;
        switch (q.kind) {
        case CATHODE_SEG_X_LOW:
        case CATHODE_SEG_R_LOW: code = GF_WLO; break;
        case CATHODE_SEG_X_HIGH:
        case CATHODE_SEG_R_HIGH: code = GF_WHI; break;
        case CATHODE_SEG_R_OPEN: code = GF_OPEN; break;
        default: code = GF_AXIS; break;
        }
        if (q.flag & CATHODE_SEG_BIAS) code |= GF_BIAS;
        if (q.flag & CATHODE_SEG_INLET) code |= GF_INLET;
        for (int t = q.lo; t <= q.hi; t++) {
            const SegPoint P = seg_point(q.kind, q.cell, t);
            std::vector<uint8_t>& fc = xk ? T.xc : T.rc;
            if (!T.act[at(P.ci, P.cj)] || T.act[at(P.gi, P.gj)])
                return fail("segment " + std::to_string(s) + " at (" + std::to_string(P.ci) + "," + std::to_string(P.cj) +
                            ") is not a plasma cell facing a non-plasma cell");
            if (fc[at(P.fi, P.fj)] != GF_NONE) return fail("segment " + std::to_string(s) + " repeats a face");
            fc[at(P.fi, P.fj)] = code;
            if (q.kind == CATHODE_SEG_R_AXIS) continue;
            const int dir = q.kind == CATHODE_SEG_R_OPEN ? 0 : 1;
            const int bias = (q.flag & CATHODE_SEG_BIAS) ? 1 : 0;
            T.gs[at(P.gi, P.gj)] = (int32_t)((at(P.ci, P.cj) << 2) | (dir << 1) | bias);
        }
        if (q.kind == CATHODE_SEG_X_HIGH && (q.flag & CATHODE_SEG_BIAS)) {
            T.anode.push_back(q.cell + 1);
            T.anode.push_back(q.lo);
            T.anode.push_back(q.hi);
        }
    }
    // every face of a plasma cell is inner or a boundary of its side
    for (int i = 1; i <= ni; i++)
        for (int j = 1; j <= nj; j++) {
            if (!T.act[at(i, j)]) continue;
            const int L = T.xc[at(i, j)] & GF_TYPE, R = T.xc[at(i + 1, j)] & GF_TYPE;
            const int S = T.rc[at(i, j)] & GF_TYPE, N = T.rc[at(i, j + 1)] & GF_TYPE;
            const bool ok = (L == GF_INNER || L == GF_WLO) && (R == GF_INNER || R == GF_WHI) &&
                            (S == GF_INNER || S == GF_WLO || S == GF_AXIS) &&
                            (N == GF_INNER || N == GF_WHI || N == GF_OPEN);
            if (!ok)
                return fail("plasma cell (" + std::to_string(i) + "," + std::to_string(j) +
                            ") has a face that is neither inner nor a boundary segment of its side");
        }
    // ghost writes after the electron and energy routines (last writer wins)
    auto ghost_list = [&](const cathode_ndd_segment* list, int cnt, std::vector<uint8_t>& bits,
                          const char* what) -> bool {
        std::vector<int32_t> owner(n, -1);  // target -> (source << 3) | (side << 1) | copy
        for (int s = 0; s < cnt; s++) {
            const cathode_ndd_segment& q = list[s];
            if (q.kind < CATHODE_SEG_X_LOW || q.kind > CATHODE_SEG_R_HIGH || q.lo > q.hi) {
                err = std::string("ndd_set_geometry: ") + what + " list: bad kind or range";
                return false;
            }
            const bool xk = q.kind == CATHODE_SEG_X_LOW || q.kind == CATHODE_SEG_X_HIGH;
            if (xk ? (q.cell < 1 || q.cell > ni || q.lo < 1 || q.hi > nj)
                   : (q.cell < 1 || q.cell > nj || q.lo < 1 || q.hi > ni)) {
                err = std::string("ndd_set_geometry: ") + what + " list outside the grid";
                return false;
            }
            for (int t = q.lo; t <= q.hi; t++) {
                const SegPoint P = seg_point(q.kind, q.cell, t);
                if (!T.act[at(P.ci, P.cj)] || T.act[at(P.gi, P.gj)]) {
                    err = std::string("ndd_set_geometry: ") + what + " list names a ghost that is not next to plasma";
                    return false;
                }
                owner[at(P.gi, P.gj)] = (int32_t)((at(P.ci, P.cj) << 3) | (side_bit(q.kind) << 1) | (q.flag ? 1 : 0));
            }
        }
        for (size_t k = 0; k < n; k++) {
            if (owner[k] < 0) continue;
            const size_t src = (size_t)owner[k] >> 3;
            const int d = (owner[k] >> 1) & 3, copy = owner[k] & 1;
            bits[src] |= (uint8_t)((1 << d) | (copy << (4 + d)));
        }
        return true;
    };
    if (!ghost_list(g.ne, g.n_ne, T.ne, "ne")) return false;
    if (!ghost_list(g.te, g.n_te, T.te, "te")) return false;
    return true;
}

}  // namespace cathode

using namespace cathode;

// initialModule.hpp:239-343: the six block boxes from the block corners
extern "C" int cathode_ndd_reference_geometry(const double* xs, const double* rs, double dx, double dr,
                                              cathode_ndd_geometry* g, int* cells6, int* xfaces6, int* rfaces6)
{
    if (!xs || !rs || !g || !(dx > 0) || !(dr > 0)) {
        set_error("ndd_reference_geometry: invalid arguments");
        return CATHODE_ERR_ARG;
    }
    const double x1 = xs[0], x2 = xs[1], x3 = xs[2], x4 = xs[3], x5 = xs[4], x6 = xs[5], xR = xs[6];
    const double r1 = rs[0], r2 = rs[1], r3 = rs[2], r4 = rs[3], r5 = rs[4], rmax = rs[5];
    auto lo = [](double v, double h) { return int(v / h + 0.5) + 1; };  // first cell at or after v
    auto hi = [](double v, double h) { return int(v / h - 0.5) + 1; };  // last cell before v
    // {i0, i1, j0, j1} per block; cells, x faces, r faces
    const int C[6][4] = {{lo(x6, dx), hi(x2, dx), lo(r3, dr), hi(r4, dr)},
                         {lo(x2, dx), hi(x4, dx), lo(r5, dr), hi(r4, dr)},
                         {lo(x2, dx), hi(x3, dx), lo(r1, dr), hi(r5, dr)},
                         {lo(x3, dx), hi(x5, dx), lo(0.0, dr), hi(r5, dr)},
                         {lo(x5, dx), hi(xR, dx), lo(0.0, dr), hi(rmax, dr)},
                         {lo(x1, dx), hi(x2, dx), lo(r1, dr), hi(r2, dr)}};
    const int X[6][4] = {{lo(x6, dx) + 1, hi(x2, dx) + 1, lo(r3, dr), hi(r4, dr)},
                         {lo(x2, dx) + 1, hi(x4, dx), lo(r5, dr), hi(r4, dr)},
                         {lo(x2, dx) + 1, hi(x3, dx) + 1, lo(r1, dr), hi(r5, dr)},
                         {lo(x3, dx) + 1, hi(x5, dx) + 1, lo(0.0, dr), hi(r5, dr)},
                         {lo(x5, dx) + 1, hi(xR, dx), lo(0.0, dr), hi(rmax, dr)},
                         {lo(x1, dx) + 1, hi(x2, dx) + 1, lo(r1, dr), hi(r2, dr)}};
    const int R[6][4] = {{lo(x6, dx), hi(x2, dx), lo(r3, dr) + 1, hi(r4, dr)},
                         {lo(x2, dx), hi(x4, dx), lo(r5, dr), hi(r4, dr)},
                         {lo(x2, dx), hi(x3, dx), lo(r1, dr) + 1, hi(r5, dr)},
                         {lo(x3, dx), hi(x5, dx), lo(0.0, dr) + 1, hi(r5, dr)},
                         {lo(x5, dx), hi(xR, dx), lo(0.0, dr) + 1, hi(rmax, dr)},
                         {lo(x1, dx), hi(x2, dx), lo(r1, dr) + 1, hi(r2, dr)}};
    for (int b = 0; b < 6; b++)
        for (int q = 0; q < 4; q++) {
            if (cells6) cells6[4 * b + q] = C[b][q];
            if (xfaces6) xfaces6[4 * b + q] = X[b][q];
            if (rfaces6) rfaces6[4 * b + q] = R[b][q];
        }
    std::memset(g, 0, sizeof(*g));
    g->n_bl = 5;
    for (int b = 0; b < 5; b++)
        for (int q = 0; q < 4; q++) {
            g->cell_box[4 * b + q] = C[b][q];
            g->xface_box[4 * b + q] = X[b][q];
            g->rface_box[4 * b + q] = R[b][q];
        }
    enum { I0 = 0, I1 = 1, J0 = 2, J1 = 3 };
    // the PC routines' boundary loops (fluidModule.hpp:2152-2249), in order
    const cathode_ndd_segment seg[] = {
        {CATHODE_SEG_X_LOW, C[0][I0], C[0][J0], C[0][J1], 0},       // z0
        {CATHODE_SEG_X_LOW, C[2][I0], C[2][J0], C[0][J0] - 1, 0},   // z1
        {CATHODE_SEG_X_LOW, C[3][I0], C[3][J0], C[2][J0] - 1, 0},   // z2
        {CATHODE_SEG_X_LOW, C[4][I0], C[1][J0], C[4][J1], 0},       // z4
        {CATHODE_SEG_X_HIGH, C[1][I1], C[1][J0], C[1][J1], 0},      // z3
        {CATHODE_SEG_X_HIGH, C[4][I1], C[4][J0], C[4][J1], CATHODE_SEG_BIAS},  // z5, the electrode at V_bias
        {CATHODE_SEG_R_LOW, C[0][J0], C[0][I0], C[0][I1], 0},       // x0
        {CATHODE_SEG_R_LOW, C[2][J0], C[2][I0], C[2][I1], 0},       // x2
        {CATHODE_SEG_R_HIGH, C[0][J1], C[0][I0], C[1][I1], CATHODE_SEG_INLET | CATHODE_SEG_UIX_BELOW},  // x1
        {CATHODE_SEG_R_HIGH, C[3][J1], C[1][I1] + 1, C[3][I1], 0},  // x4
        {CATHODE_SEG_R_OPEN, C[4][J1], C[4][I0], C[4][I1], 0},      // x5
        {CATHODE_SEG_R_AXIS, C[3][J0], C[3][I0], C[4][I1], 0},      // x6, the axis row of blocks 3 and 4
    };
    g->n_seg = (int)(sizeof(seg) / sizeof(seg[0]));
    for (int s = 0; s < g->n_seg; s++) g->seg[s] = seg[s];
    // ghosts after the electron routine (:3873-3913): the walls of blocks 0
    // and 1 get n_e = 0, their open sides a copy
    const cathode_ndd_segment ne[] = {
        {CATHODE_SEG_X_LOW, C[0][I0], C[0][J0], C[0][J1], 0},      // chamber
        {CATHODE_SEG_X_LOW, C[1][I0], C[1][J0], C[0][J0] - 1, 0},  // magnet
        {CATHODE_SEG_R_LOW, C[0][J0], C[0][I0], C[0][I1], 0},      // antenna
        {CATHODE_SEG_X_HIGH, C[1][I1], C[1][J0], C[1][J1], 1},     // right open
        {CATHODE_SEG_R_HIGH, C[0][J1], C[0][I0], C[1][I1], 1},     // upper open
    };
    g->n_ne = 5;
    g->n_te = 5;
    for (int s = 0; s < 5; s++) {
        g->ne[s] = ne[s];
        g->te[s] = ne[s];
        g->te[s].flag = 1;  // Te ghosts copy the cell on every side (:4906-4946)
    }
    return CATHODE_OK;
}
