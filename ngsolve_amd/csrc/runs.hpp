// Pattern-runs layout of a CSR matrix, shared by the single and the batched
// SpMV: up to kMaxRun consecutive rows with one column list store that list
// once and their values column by column ([L][R]); runs are grouped into
// blocks of equal run length with at most max_vals values and max_vals rows.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "csr.hpp"

namespace ngsamd {

// Exact k / R for 0 <= k <= 4096 and 1 <= R <= 8 as (k * ceil(2^16 / R)) >> 16:
// the rounding error k (ceil(2^16/R) - 2^16/R) / 2^16 stays below 1/16 while
// the fractional part of k / R is at most 7/8.
__device__ __forceinline__ int RunDiv(int k, int inv) { return (k * inv) >> 16; }

// Splits the rows into pattern runs (at most kMaxRun rows, at most max_vals
// values per run unless the run is one row) and the runs into blocks of equal
// run length with at most max_vals values and max_vals rows (or one longer
// row).  Fills the
// host arrays of the runs layout; returns the number of shared columns.
// With v null only the structure is built (rvals stays empty; val_ptr still
// counts the values), for values filled in on the device.
// With seg (nseg + 1 row offsets of independent problems), runs and blocks
// never cross a problem boundary and blk_seg receives each block's problem.
inline int64_t BuildRuns(int64_t h, const int64_t* rp, const int* c, const double* v, int max_vals,
                         std::vector<int>& run_row, std::vector<int>& run_ptr, std::vector<int>& val_ptr,
                         std::vector<int>& rcols, std::vector<double>& rvals, std::vector<int>& blk_run,
                         std::vector<int>& blk_len, const int64_t* seg = nullptr, int nseg = 0,
                         std::vector<int>* blk_seg = nullptr) {
  run_row.assign(1, 0);
  run_ptr.assign(1, 0);
  val_ptr.assign(1, 0);
  rcols.clear();
  rvals.clear();
  blk_run.assign(1, 0);
  blk_len.clear();
  if (blk_seg) blk_seg->clear();
  int64_t acc = 0, rows = 0, nval = 0;
  int cur_len = 0;
  int p = 0, cur_seg = 0;  // problem of row r, problem of the open block
  for (int64_t r = 0; r < h;) {
    while (seg && p < nseg && r >= seg[p + 1]) p++;
    const int64_t row_end = seg && p < nseg ? seg[p + 1] : h;
    const int64_t L = rp[r + 1] - rp[r];
    int R = 1;
    const int rmax = L > 0 ? (int)std::max<int64_t>(1, std::min<int64_t>(kMaxRun, max_vals / L)) : kMaxRun;
    while (R < rmax && r + R < row_end && rp[r + R + 1] - rp[r + R] == L &&
           std::equal(c + rp[r], c + rp[r] + L, c + rp[r + R]))
      R++;
    const int64_t nv = int64_t(R) * L;
    const int u = (int)run_row.size() - 1;
    if (u > blk_run.back() && (R != cur_len || acc + nv > max_vals || rows + R > max_vals || p != cur_seg)) {
      blk_len.push_back(cur_len);
      blk_run.push_back(u);
      if (blk_seg) blk_seg->push_back(cur_seg);
      acc = 0;
      rows = 0;
    }
    cur_seg = p;
    cur_len = R;
    acc += nv;
    rows += R;
    rcols.insert(rcols.end(), c + rp[r], c + rp[r] + L);
    if (v)
      for (int64_t j = 0; j < L; j++)
        for (int q = 0; q < R; q++) rvals.push_back(v[rp[r + q] + j]);
    nval += nv;
    r += R;
    run_row.push_back((int)r);
    run_ptr.push_back((int)rcols.size());
    val_ptr.push_back((int)nval);
  }
  const int nruns = (int)run_row.size() - 1;
  if (nruns > blk_run.back()) {
    blk_len.push_back(cur_len);
    blk_run.push_back(nruns);
    if (blk_seg) blk_seg->push_back(cur_seg);
  }
  return (int64_t)rcols.size();
}


// Regroups existing runs into blocks of equal run length with at most
// max_vals values and max_vals rows (a run longer than max_vals alone), and
// never across a problem boundary (run_seg: problem of each run, or empty).
// The grouping rule of BuildRuns, for another block size on the same runs.
inline void GroupRunBlocks(const std::vector<int>& run_row, const std::vector<int>& val_ptr,
                           const std::vector<int>& run_seg, int max_vals, std::vector<int>& blk_run,
                           std::vector<int>& blk_len, std::vector<int>* blk_seg) {
  blk_run.assign(1, 0);
  blk_len.clear();
  if (blk_seg) blk_seg->clear();
  const int nruns = (int)run_row.size() - 1;
  int64_t acc = 0, rows = 0;
  int cur_len = 0, cur_seg = 0;
  for (int u = 0; u < nruns; u++) {
    const int R = run_row[u + 1] - run_row[u];
    const int64_t nv = val_ptr[u + 1] - val_ptr[u];
    const int p = run_seg.empty() ? 0 : run_seg[u];
    if (u > blk_run.back() && (R != cur_len || acc + nv > max_vals || rows + R > max_vals || p != cur_seg)) {
      blk_len.push_back(cur_len);
      blk_run.push_back(u);
      if (blk_seg) blk_seg->push_back(cur_seg);
      acc = 0;
      rows = 0;
    }
    cur_seg = p;
    cur_len = R;
    acc += nv;
    rows += R;
  }
  if (nruns > blk_run.back()) {
    blk_len.push_back(cur_len);
    blk_run.push_back(nruns);
    if (blk_seg) blk_seg->push_back(cur_seg);
  }
}

}  // namespace ngsamd
