// Host phase 2: make the triangulation exactly Delaunay when the concurrent
// GPU flipping left faces it could not flip (the reference's CPU fix-up
// chain, src/lib.rs:227-307: cpu_flip_violations -> full_bw_reconstruction
// -> star splaying).
#pragma once

#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/wtet.h"

struct RepairReport {
  uint32_t regions = 0;      // regions re-triangulated
  uint32_t max_region = 0;   // largest region (tets)
  bool needs_full = false;   // the repair reached a row the sparse view lacks
  bool too_large = false;    // a region outgrew the cap the caller set
  uint32_t fetches = 0;      // rows fetched into the sparse view on demand
  uint32_t inserted = 0;     // points inserted by insert_points
  std::string message;
};

// The mesh in the result format (input ids, super n..n+3, infinity n+4;
// adjacency (t << 5) | f) in caller-owned buffers with room for `cap` tets.
// The repair edits it in place: new tets reuse freed slots or are appended,
// holes are filled from the end; `changed` lists every slot whose tet or
// adjacency row was written (sorted, < count on return).
struct MeshBuf {
  uint32_t* tets;
  uint32_t* adj;
  size_t count;
  size_t cap;
  std::vector<uint32_t> changed;
  // sparse view: when set, only rows r < known with present[r] ==
  // present_mark hold data (rows >= known are the repair's own).  Reaching another row calls
  // fetch(row), which copies rows around it in and marks them present, and
  // retries; without fetch, or when it returns false, the repair stops with
  // needs_full.  Rows the repair has written are never fetched again: every
  // neighbour of a written row is read before the write.
  uint8_t* present = nullptr;
  uint8_t present_mark = 1;
  size_t known = 0;
  std::function<bool(uint32_t)> fetch;
  // optional: the rows around every listed row in one pass (one device
  // round trip for all the rows a round of region plans reached); returns
  // false when it cannot make them present
  std::function<bool(const std::vector<uint32_t>&)> fetch_many;
};

// xyzn: n normalized points (input id order).  viol: (t << 2) | f of
// non-Delaunay faces.  On success the mesh is the exact Delaunay
// triangulation.  region_cap > 0: stop with WTET_ERR_NOT_CONVERGED and
// rep->too_large once a region holds more than region_cap tets (the mesh is
// then partly repaired and only fit to be discarded); 0: no limit.
wtet_status repair_delaunay(const float* xyzn, uint32_t n, MeshBuf& mesh,
                            const std::vector<uint32_t>& viol, int mode, RepairReport* rep,
                            uint32_t region_cap = 0);

// Inserts the input ids `ids` (points the insertion rounds did not insert,
// src/phase1.rs:26 stops them at max_insert_iterations) into a mesh that is
// already the exact Delaunay triangulation of the others, one Bowyer-Watson
// cavity each (src/phase2.rs:383-606 inserts the same way); the mesh stays
// Delaunay.  Every row must be on the host (no sparse view).  Exact
// duplicates have no cavity and are skipped.
wtet_status insert_points(const float* xyzn, uint32_t n, MeshBuf& mesh, std::vector<uint32_t> ids,
                          int mode, RepairReport* rep);

// input ids 0..n-1 used by no tet among `count` rows
std::vector<uint32_t> missing_points(const uint32_t* tets, size_t count, uint32_t n);
