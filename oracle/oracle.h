/*
 * CPU oracle for the phase-1 Delaunay path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, on the CPU and in plain C, what the reference
 * computes on the triangulation path.  It is the checker: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it.  The product (wtet_amd) never links or calls it.
 *
 * Contents and the reference code each part follows:
 *   - exact orient3d / insphere signs           src/predicates.rs:14-21
 *     (third-party geometry-predicates 0.3.0 = Shewchuk's predicates; here
 *     re-derived with floating-point expansion arithmetic, Shewchuk 1997)
 *   - orient3d with 14-level SoS                src/predicates.rs:106-208
 *   - orient4d with 49-level SoS                src/predicates.rs:450-675
 *     (lifted sub-determinants :218-435, plain-f64 orient2d :73-75)
 *   - normalize_points (f32)                    src/lib.rs:313-337
 *   - Morton code                               src/morton.rs:22-81
 *   - initial stride + stratified order         src/gpu/mod.rs:533-600
 *   - serial Bowyer-Watson over points + 4 super vertices
 *                                               src/phase2.rs:383-606
 *     (super vertices src/gpu/mod.rs:152-162, orientation init.wgsl:98-118)
 *   - Delaunay audit (orient4d on shared faces) src/lib.rs:25-78
 *   - adjacency rebuild by face matching        src/phase2.rs:1460-1499
 *
 * Sign conventions (as in the reference):
 *   orient3d_sign(a,b,c,d) = sign det[a-d; b-d; c-d]  (>0: d below plane abc)
 *   insphere_sign(a,b,c,d,e) = sign of Shewchuk's insphere (>0: e inside
 *     when orient3d(a,b,c,d) > 0)
 *   orient3d_sos returns +1 when the (perturbed) raw determinant is < 0
 *   ("OrientPos" in gDel3D), -1 otherwise; never 0.
 *   orient4d returns +1 (outside) / -1 (inside) for a tet stored with raw
 *   orient3d < 0; never 0.
 */
#ifndef WTET_ORACLE_H
#define WTET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- exact predicates on f64 inputs ------------------------------------ */
int o_orient3d_sign(const double *a, const double *b, const double *c,
                    const double *d);
int o_insphere_sign(const double *a, const double *b, const double *c,
                    const double *d, const double *e);
/* exact sign of det [[p0 p1 |p|^2 1]] over 4 points (coords as given) */
int o_orient3d_lifted_sign(const double *a, const double *b, const double *c,
                           const double *d);
/* exact sign of det [[p0 |p|^2 1]] over 3 points */
int o_orient2d_lifted_sign(const double *a, const double *b, const double *c);
/* exact sign of |a|^2 - |b|^2 */
int o_orient1d_lifted_sign(const double *a, const double *b);
/* plain f64 (a0-c0)*(b1-c1) - (a1-c1)*(b0-c0), IEEE round-to-nearest,
 * no fused multiply-add -- replicates src/predicates.rs:73-75 bit for bit */
double o_orient2d_plain(const double *a, const double *b, const double *c);

/* ---- SoS predicates (reference semantics) ------------------------------ */
int o_orient3d_sos(const double *a, const double *b, const double *c,
                   const double *d, uint32_t ia, uint32_t ib, uint32_t ic,
                   uint32_t id);
int o_orient4d(const double *a, const double *b, const double *c,
               const double *d, const double *e, uint32_t ia, uint32_t ib,
               uint32_t ic, uint32_t id, uint32_t ie);
/* the cascade alone (used when the exact insphere is zero); also reports the
 * depth that decided (1..49) through *depth when non-null */
int o_orient4d_sos_only(const double *a, const double *b, const double *c,
                        const double *d, const double *e, uint32_t ia,
                        uint32_t ib, uint32_t ic, uint32_t id, uint32_t ie,
                        int *depth);
int o_orient3d_sos_depth(const double *a, const double *b, const double *c,
                         const double *d, uint32_t ia, uint32_t ib,
                         uint32_t ic, uint32_t id, int *depth);

/* batched helpers for the Python side: n queries, points as f64 triples */
void o_orient3d_sign_batch(const double *pts, int64_t n, int32_t *out);
void o_insphere_sign_batch(const double *pts, int64_t n, int32_t *out);
void o_orient3d_sos_batch(const double *pts, const uint32_t *idx, int64_t n,
                          int32_t *out, int32_t *depth);
void o_orient4d_batch(const double *pts, const uint32_t *idx, int64_t n,
                      int32_t *out, int32_t *depth);

/* ---- point-set helpers ------------------------------------------------- */
void o_normalize_points(const float *xyz, int64_t n, float *out,
                        float *bbox_min3, float *scale);
uint32_t o_morton_code(const float *p, const float *mn, const float *mx);
uint32_t o_expand_bits(uint32_t v);
int64_t o_initial_stride(int64_t n);
void o_stratified_order(int64_t n, int64_t stride, int64_t *order);

/* ---- serial Bowyer-Watson ---------------------------------------------- */
/* Triangulates the n normalized points (f32) together with the 4 super
 * vertices at ids n..n+3.  SoS ranks are the point ids.  Returns the number
 * of tets written to *tets_out (malloc'd, 4*count u32, free with o_free),
 * or -1 on failure.  Every tet has raw orient3d <= 0 and orient3d_sos = +1.
 * stats (may be null) receives [seed_scans, orient_inconsistencies,
 * exact_calls, filter_calls]. */
int64_t o_delaunay_bw(const float *pts, int64_t n, uint32_t **tets_out,
                      int64_t *stats);
void o_free(void *p);

/* full_bw_reconstruction as written (src/phase2.rs:383-606), the
 * reference's CPU path: points inserted in input order, each seeded by a
 * linear scan over every tet created so far (dead ones skipped) for the
 * first whose circumsphere holds it, BFS cavity visited in ascending tet
 * order, new tets appended (never reused), new-tet faces linked by pairwise
 * search.  Same output contract as o_delaunay_bw.  O(N^2): the CPU
 * baseline on small inputs only. */
int64_t o_delaunay_bw_reference(const float *pts, int64_t n, uint32_t **tets_out);
/* Many independent normalized clouds on `threads` OpenMP threads, one
 * o_delaunay_bw each: counts[c] = tets of cloud c; returns the total or -1. */
int64_t o_delaunay_bw_many(const float *pts, const uint64_t *off, int64_t C, int threads,
                           int64_t *counts);

/* Keep tets whose 4 ids are all < n, sort each quad, sort the list.
 * Writes into out (capacity m*4), returns count. */
int64_t o_canonical_real(const uint32_t *tets, int64_t m, uint32_t n,
                         uint32_t *out);

/* Rebuild adjacency by face matching: adj[t*4+f] = (t' << 5) | f' or
 * 0xFFFFFFFF; face f of a tet is the face opposite vertex f. */
void o_rebuild_adjacency(const uint32_t *tets, int64_t m, uint32_t *adj);

/* Count interior shared faces whose orient4d says "inside" (reference
 * check_delaunay_quality).  pts: n normalized f32 points (super vertices
 * appended internally).  only_real != 0 -> skip tets touching ids >= n. */
int64_t o_count_violations(const float *pts, int64_t n,
                           const uint32_t *tets, const uint32_t *adj,
                           int64_t m, int only_real);

#ifdef __cplusplus
}
#endif
#endif
