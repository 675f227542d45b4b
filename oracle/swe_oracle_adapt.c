/*
 * swe_oracle_adapt.c -- CPU restatement of the sam(oa)^2 SWE remeshing step.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP remesher.
 *
 * One adaption traversal of the reference (SWE.f90:387-390) is
 *   1. the conformity check (Conformity/Conformity.f90:172-208): edge flags
 *      `refine` and `coarsen` set from the cells' refinement requests by
 *      mark_edges (:1244-1268) and closed under mark_edges_integrity
 *      (:1271-1317), traversal after traversal until nothing changes;
 *   2. the adaptive traversal (SFC_generic_adaptive_traversal.f90) that splits
 *      every cell whose hypotenuse is marked (a marked leg splits that child
 *      again), merges every pair of sibling cells whose legs are marked for
 *      coarsening, and copies the others;
 *   3. the state transfer of SWE_adapt.f90:95-169: a split copies the state
 *      (order-0 t_basis_Q_split, Tools_lagrange_basis.f90:178-186), a merge
 *      averages it, 0.5 (Q1 + Q2) (merge_vector, :139-144); b of a split cell
 *      is the parent's here (the reference evaluates the scenario's
 *      bathymetry, which is the parent's for a flat bed).
 *
 * Mesh conventions as swe_oracle.h: cells in forward curve order, corners in
 * the order (entry, right angle, exit), edge k opposite corner k, edge 1 the
 * hypotenuse.  A cell (e, r, x) splits into (e, m, r) and (r, m, x), m the
 * midpoint of the hypotenuse (SFC_node_traversal.f90:138-143), with plotter
 * types child(p) and -child(-p) (:257,270).
 *
 * The new neighbour table is rebuilt from the new corners (every edge of a
 * conforming mesh is named by its midpoint), independently of the way the
 * HIP remesher derives it from the old table.
 *
 * Parity pins: tests/test_remesh.py runs this function over the outputs of
 * the reference binary with remeshing after every step (-nadapt 1) and
 * requires the reference's next mesh (corners, depths, plotter types) and,
 * after one oracle step, its next state, bit for bit.
 */
#include "swe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static const int8_t child_type[17] = {3, 2, 1, 8, 7, 6, 5, 4, 0,
                                      -6, -7, -8, -1, -2, -3, -4, -5};

typedef struct {
    int64_t kx, ky;
    int64_t slot;
} mid_key;

static int cmp_mid(const void *a, const void *b) {
    const mid_key *x = (const mid_key *)a, *y = (const mid_key *)b;
    if (x->kx != y->kx) return x->kx < y->kx ? -1 : 1;
    if (x->ky != y->ky) return x->ky < y->ky ? -1 : 1;
    return x->slot < y->slot ? -1 : (x->slot > y->slot);
}

/* Conformity.f90:1271-1317 on the three edges of one cell; edge ids h (the
 * hypotenuse), l, r (the legs).  Returns nonzero if a flag changed. */
static int integrity(int64_t h, int64_t l, int64_t r, int removable, uint8_t *ref,
                     uint8_t *coa) {
    const uint8_t rh = ref[h] | ref[l] | ref[r];
    const uint8_t c = (uint8_t)(!rh && coa[l] && coa[r] && removable);
    const int changed = rh != ref[h] || c != coa[l] || c != coa[r];
    ref[h] = rh;
    coa[l] = c;
    coa[r] = c;
    return changed;
}

int64_t swo_adapt(int64_t n, const int8_t *depth, const int8_t *plotter, const int32_t *nbr,
                  const double *coords, const int8_t *flags, int dmin, int dmax,
                  const double *eta, const double *hu, const double *hv, const double *b,
                  int64_t cap, int8_t *depth_out, int8_t *plotter_out, int32_t *nbr_out,
                  double *coords_out, double *eta_out, double *hu_out, double *hv_out,
                  double *b_out, int *sweeps) {
    if (n <= 0) return -1;
    /* shared edge ids: one per interior edge, one per boundary edge */
    int64_t *eid = (int64_t *)malloc((size_t)(3 * n) * sizeof(int64_t));
    if (!eid) return -2;
    for (int64_t i = 0; i < 3 * n; ++i) eid[i] = -1;
    int64_t ne = 0;
    for (int64_t c = 0; c < n; ++c)
        for (int k = 0; k < 3; ++k) {
            if (eid[3 * c + k] >= 0) continue;
            eid[3 * c + k] = ne;
            const int32_t o = nbr[3 * c + k];
            if (o >= 0) {
                int ko = -1;
                for (int j = 0; j < 3; ++j)
                    if (nbr[3 * (int64_t)o + j] == c) ko = j;
                if (ko < 0) {
                    free(eid);
                    return -3;
                }
                eid[3 * (int64_t)o + ko] = ne;
            }
            ++ne;
        }
    uint8_t *ref = (uint8_t *)malloc((size_t)ne), *coa = (uint8_t *)malloc((size_t)ne);
    if (!ref || !coa) {
        free(eid);
        free(ref);
        free(coa);
        return -2;
    }
    /* initial conformity traversal (:668-773): every edge starts as
     * refine = false, coarsen = true, then mark_edges per cell */
    memset(ref, 0, (size_t)ne);
    memset(coa, 1, (size_t)ne);
    for (int64_t c = 0; c < n; ++c) {
        const int64_t h = eid[3 * c + 1], l = eid[3 * c + 0], r = eid[3 * c + 2];
        int f = flags[c];
        if (f > 0 && depth[c] >= dmax) f = 0; /* volume_op :339 guards */
        if (f < 0 && depth[c] <= dmin) f = 0; /* :342 */
        coa[h] = 0;                           /* the hypotenuse is never coarsened */
        if (f == 0) coa[l] = 0;
        if (f == 1) ref[h] = 1;
        integrity(h, l, r, depth[c] > 0, ref, coa);
    }
    /* update conformity traversals (:995-1084) until the grid is conform */
    int it = 1, changed = 1;
    while (changed) {
        changed = 0;
        for (int64_t c = 0; c < n; ++c)
            changed |= integrity(eid[3 * c + 1], eid[3 * c + 0], eid[3 * c + 2], depth[c] > 0,
                                 ref, coa);
        ++it;
    }
    if (sweeps) *sweeps = it;

    /* tree position: leaf c covers [off, off + 2^(60 - depth)) of the two
     * roots [0, 2^61); a first child starts at a multiple of its parent's size */
    int64_t off = 0, m = 0, status = 0;
    for (int64_t c = 0; c < n && status == 0; ++c) {
        const int d = depth[c];
        const int64_t size = (int64_t)1 << (60 - d);
        const int64_t e = eid[3 * c + 1];
        const double *x = coords ? coords + 6 * c : NULL;
        if (ref[e]) {
            /* split: hypotenuse, then the children whose hypotenuse (a leg of
             * this cell) is marked; leg 2 belongs to the first child, leg 0 to
             * the second */
            const int p1 = child_type[plotter[c] + 8], p2 = -child_type[-plotter[c] + 8];
            const int split[2] = {ref[eid[3 * c + 2]], ref[eid[3 * c + 0]]};
            double kid[2][6];
            if (x) {
                const double mx = 0.5 * (x[0] + x[4]), my = 0.5 * (x[1] + x[5]);
                const double k1[6] = {x[0], x[1], mx, my, x[2], x[3]};
                const double k2[6] = {x[2], x[3], mx, my, x[4], x[5]};
                memcpy(kid[0], k1, sizeof k1);
                memcpy(kid[1], k2, sizeof k2);
            }
            const int pk[2] = {p1, p2};
            for (int s = 0; s < 2; ++s) {
                const int nk = split[s] ? 2 : 1;
                if (m + nk > cap) {
                    status = -4;
                    break;
                }
                for (int j = 0; j < nk; ++j) {
                    int pt = pk[s], dd = d + 1;
                    double g[6];
                    if (x) memcpy(g, kid[s], sizeof g);
                    if (split[s]) {
                        dd = d + 2;
                        pt = j == 0 ? child_type[pk[s] + 8] : -child_type[-pk[s] + 8];
                        if (x) {
                            const double *y = kid[s];
                            const double mx = 0.5 * (y[0] + y[4]), my = 0.5 * (y[1] + y[5]);
                            const double a1[6] = {y[0], y[1], mx, my, y[2], y[3]};
                            const double a2[6] = {y[2], y[3], mx, my, y[4], y[5]};
                            memcpy(g, j == 0 ? a1 : a2, sizeof g);
                        }
                    }
                    depth_out[m] = (int8_t)dd;
                    plotter_out[m] = (int8_t)pt;
                    if (coords_out && x) memcpy(coords_out + 6 * m, g, sizeof g);
                    eta_out[m] = eta[c];
                    hu_out[m] = hu[c];
                    hv_out[m] = hv[c];
                    b_out[m] = b[c];
                    ++m;
                }
            }
        } else if (coa[eid[3 * c + 0]]) {
            /* merge with the sibling: only a first child whose sibling is the
             * next leaf gets here with both legs marked */
            const int64_t parent_size = (int64_t)1 << (61 - d);
            const int first = (off % parent_size) == 0;
            if (!first || c + 1 >= n || depth[c + 1] != d || !coa[eid[3 * (c + 1)]] ||
                ref[eid[3 * (c + 1) + 1]]) {
                status = -5;
                break;
            }
            if (m + 1 > cap) {
                status = -4;
                break;
            }
            int parent = 0;
            for (int q = -8; q <= 8; ++q)
                if (q != 0 && child_type[q + 8] == plotter[c]) parent = q;
            depth_out[m] = (int8_t)(d - 1);
            plotter_out[m] = (int8_t)parent;
            if (coords_out && x) {
                const double *y = coords + 6 * (c + 1);
                const double g[6] = {x[0], x[1], x[4], x[5], y[4], y[5]};
                memcpy(coords_out + 6 * m, g, sizeof g);
            }
            eta_out[m] = 0.5 * (eta[c] + eta[c + 1]);
            hu_out[m] = 0.5 * (hu[c] + hu[c + 1]);
            hv_out[m] = 0.5 * (hv[c] + hv[c + 1]);
            b_out[m] = 0.5 * (b[c] + b[c + 1]);
            ++m;
            off += size;
            ++c; /* the sibling is consumed */
        } else {
            if (m + 1 > cap) {
                status = -4;
                break;
            }
            depth_out[m] = depth[c];
            plotter_out[m] = plotter[c];
            if (coords_out && x) memcpy(coords_out + 6 * m, x, 6 * sizeof(double));
            eta_out[m] = eta[c];
            hu_out[m] = hu[c];
            hv_out[m] = hv[c];
            b_out[m] = b[c];
            ++m;
        }
        off += size;
    }
    free(eid);
    free(ref);
    free(coa);
    if (status) return status;

    if (coords && coords_out && nbr_out) {
        /* neighbours of the new mesh from its corners: edge midpoints are
         * dyadic with at most 31 fractional bits here (depth <= 60) */
        mid_key *keys = (mid_key *)malloc((size_t)(3 * m) * sizeof(mid_key));
        if (!keys) return -2;
        const double scale = ldexp(1.0, 33);
        for (int64_t c = 0; c < m; ++c) {
            const double *x = coords_out + 6 * c;
            for (int k = 0; k < 3; ++k) {
                const int a = (k + 1) % 3, bb = (k + 2) % 3;
                keys[3 * c + k].kx = (int64_t)llround(0.5 * (x[2 * a] + x[2 * bb]) * scale);
                keys[3 * c + k].ky = (int64_t)llround(0.5 * (x[2 * a + 1] + x[2 * bb + 1]) * scale);
                keys[3 * c + k].slot = 3 * c + k;
            }
        }
        qsort(keys, (size_t)(3 * m), sizeof(mid_key), cmp_mid);
        for (int64_t i = 0; i < 3 * m; ++i) nbr_out[i] = -1;
        for (int64_t i = 0; i + 1 < 3 * m; ++i)
            if (keys[i].kx == keys[i + 1].kx && keys[i].ky == keys[i + 1].ky) {
                const int64_t s0 = keys[i].slot, s1 = keys[i + 1].slot;
                nbr_out[s0] = (int32_t)(s1 / 3);
                nbr_out[s1] = (int32_t)(s0 / 3);
                ++i;
            }
        free(keys);
    }
    return m;
}
