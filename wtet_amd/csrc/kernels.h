// Launchers for the triangulation kernels (defined in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prep.h"

namespace wtet {

constexpr int kBlock = 256;  // 4 wavefronts of 64

inline uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// per-cloud bounding box only (the first kernels of launch_prepare): mm
// gets 6 order-preserving words per cloud (min xyz, max xyz)
void launch_minmax(const float* xyz, uint32_t n, const Clouds& K, uint32_t* mm, hipStream_t s);
// point preparation for one or many clouds (layout: prep.h):
// per-cloud normalization (src/lib.rs:313-337) and (cloud, Morton) keys
void launch_prepare(const float* xyz, uint32_t n, const Clouds& K, uint32_t* mm, float* xyzn,
                    uint64_t* keys, uint32_t* ids, hipStream_t s);
// sorted points -> vertices, plus every cloud's super vertices and infinity
void launch_gather_points(const float* xyzn, const uint32_t* ids, const uint64_t* keys,
                          uint32_t n, const Clouds& K, float4* pts, hipStream_t s);
// five start tets per cloud in slots 5c..5c+4
void launch_start_tets(const Mesh& M, const Clouds& K, hipStream_t s);
// stable radix sort of (key, id) pairs (sort.h): keys/ids in, keys_tmp/ids_tmp out
hipError_t sort_pairs(uint64_t* keys, uint32_t* ids, uint64_t* keys_tmp, uint32_t* ids_tmp,
                      uint32_t n, int end_bit, void* temp, size_t temp_bytes, hipStream_t s);
size_t sort_temp_bytes(uint32_t n, int end_bit);
size_t sort_temp_bytes_u32(uint32_t n, int end_bit);
// bits of an output sort key: clouds 0..C-1 and C for free slots
inline int cloud_key_bits(uint32_t C) {
  int bits = 1;
  while ((1ull << bits) <= C) ++bits;
  return bits;
}
// uninserted list = every point except redundant exact duplicates;
// ctr[C_NPTS] = its length
void launch_init_uninserted(const float4* pts, const uint64_t* keys, uint32_t n, const Clouds& K,
                            uint32_t* blk, uint32_t* uvert, uint32_t* utet, uint32_t* ctr,
                            hipStream_t s);

// Hilbert mode: reorder the uninserted list into the reference's
// stratified order, per cloud (scratch: key, key_tmp, val_tmp of U entries)
hipError_t launch_stratify(uint32_t* uvert, uint32_t* utet, uint32_t U, const Clouds& K,
                           uint64_t* key, uint64_t* key_tmp, uint32_t* val_tmp, void* temp,
                           size_t temp_bytes, hipStream_t s);

// rounds.  The round kernels read their counts and the round number from
// ctr (C_ULEN, C_WIN, C_SURV, C_ROUND); the count arguments only bound the
// launch grids, a power of two of workgroups (round_grid), so a captured
// round replays for any count up to its bound.
uint32_t round_grid(uint64_t bound);
void launch_vote(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                 hipStream_t s);
// ring: insert only where no tet within `ring` face-steps has a smaller vote
// (blk: round_grid(U) entries)
void launch_pick(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                 uint2* wlist, uint32_t* uvert2, uint32_t* utet2, uint32_t* blk, int ring,
                 hipStream_t s);
// the two active lists (current = act[ctr[C_PAR]]), flip proposals and the
// executed flips of one iteration
struct FlipLists {
  uint32_t* act[2];
  uint4* prop;
  uint4* out;    // two uint4 per executed flip (flip_exec_body)
  uint32_t* slow;  // entries the FP32-only check deferred (ctr[C_SLOW_T])
};
// appends the children to the current active list after its ctr[C_A]
// entries and updates ctr[C_A]
void launch_split(const Mesh& M, const uint2* wlist, uint32_t W, const FlipLists& L, hipStream_t s);
// ctr[C_ULEN] = ctr[C_SURV], ctr[C_ROUND] += 1: the next round, on the device
void launch_round_next(uint32_t* ctr, hipStream_t s);
// filtered mode: an FP32-only kernel, then the full predicates on the points
// it deferred (slow, U entries of scratch); exact mode: one full kernel
void launch_split_points(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                         uint64_t* st, uint32_t* slow, hipStream_t s);
// one flip iteration = begin, check, decide+execute, update, end; counts,
// tag, round and parity are read from ctr (device-driven, graph-capturable);
// grid = persistent blocks of the three list kernels
void launch_iter_begin(uint32_t* ctr, hipStream_t s);
// filtered mode: the FP32-only check (deferring undecided tets), then the
// full check of the deferred tets; exact mode: the full check, nothing deferred
void launch_flip_check(const Mesh& M, const FlipLists& L, uint32_t grid, uint64_t* st,
                       hipStream_t s);
void launch_flip_check_deferred(const Mesh& M, const FlipLists& L, uint32_t grid, uint64_t* st,
                                hipStream_t s);
void launch_flip_decide_exec(const Mesh& M, const FlipLists& L, uint32_t grid, hipStream_t s);
void launch_flip_update(const Mesh& M, const FlipLists& L, uint32_t grid, hipStream_t s);
// up to k iterations of a list of at most `limit` entries in one workgroup
// (no deferral: the full predicates inline); sums as launch_iter_end
void launch_flip_small(const Mesh& M, const FlipLists& L, uint64_t* st, uint64_t* sums, int k,
                       uint32_t limit, hipStream_t s);
// sums[0] += checked tets, sums[1] += flips, sums[2] += iterations
void launch_iter_end(uint32_t* ctr, uint64_t* sums, hipStream_t s);
// end of one iteration followed by the begin of the next
void launch_iter_step(uint32_t* ctr, uint64_t* sums, hipStream_t s);
void launch_relocate(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                     uint64_t* st, uint32_t* slow, hipStream_t s);
// filtered mode: an FP32-only pass, then the full predicates on the tets it
// could not decide (slow: T entries of scratch); exact mode: one full pass
void launch_audit(const Mesh& M, uint32_t T, uint32_t* rec, uint32_t cap, uint64_t* st,
                  uint32_t* slow, hipStream_t s);

// output: live tets grouped by cloud (rows of cloud c from cstart[c],
// cstart[C] = ctr[C_ALIVE] = rows), cloud-local vertex ids and
// (row << 5) | face adjacency; slot2pos maps slots to rows
struct OutputBufs {
  uint32_t* blk;
  uint32_t* olist;     // T
  uint32_t* okey;      // T (many clouds)
  uint32_t* olist2;    // T (many clouds)
  uint32_t* okey2;     // T (many clouds)
  uint32_t* slot2pos;  // T
  uint32_t* cstart;    // C + 1
  void* temp;          // sort temporary (many clouds)
  size_t temp_bytes;
  uint32_t* out_tets;  // rows * 4 (null: only slot2pos / cstart)
  uint32_t* out_adj;
  uint32_t* vrank;     // scratch, one word per vertex (K.V + 1): each vertex's SoS rank
};
hipError_t launch_output(const Mesh& M, uint32_t T, const Clouds& K, const OutputBufs& B, hipStream_t s);

// violation records (slot << 2 | face) -> output indices
void launch_map_records(uint32_t* rec, uint32_t n, const uint32_t* slot2idx, hipStream_t s);
// rows within `rings` face-steps of the recorded faces' rows and of the last
// `tail` rows: list[0, *cnt) (mark: M words of scratch, bounds: rings + 2).
// more=true keeps the rows already listed and appends those around the new
// seeds that are not (list[old *cnt, new *cnt)).
void launch_neighbourhood(const uint32_t* rec, uint32_t nrec, uint32_t M, uint32_t tail, int rings,
                          const uint32_t* adj, uint32_t* mark, uint32_t* list, uint32_t* cnt,
                          uint32_t* bounds, bool more, hipStream_t s);
// packed (row, tet[4], adj[4]) records of the listed rows
void launch_nbr_gather(const uint32_t* list, uint32_t n, const uint32_t* tets, const uint32_t* adj,
                       uint32_t* out, hipStream_t s);
// write (slot, tet[4], adj[4]) records into the output arrays
void launch_apply_patch(const uint32_t* patch, uint32_t n, uint32_t* tets, uint32_t* adj,
                        hipStream_t s);

// counters
void launch_set_ctr(uint32_t* ctr, int idx, uint32_t v, int idx2, uint32_t v2, hipStream_t s);

// fills
void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
void launch_fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s);
// plain streaming copy (16 B per lane), also the HBM-counter calibration kernel
void launch_copy16(uint4* dst, const uint4* src, uint64_t n, hipStream_t s);

// predicate batches (for checking the device predicates)
void launch_pred_batch(int kind, const float* pts, const uint32_t* idx, uint32_t n, int mode,
                       int32_t* out, uint64_t* st, hipStream_t s);

}  // namespace wtet
