"""CPU check of the LDS-DMA plans of the pipelined Cartesian kernel
(dg_operator_pipe.hpp), for every (dim, degree) it is compiled for.  A host
program evaluates each PipeShape's compile-time plan and checks that:
  * every instruction reads inside its source piece (the cell's block, the
    accumulator's, one neighbour field, the coefficient block) and writes
    inside that piece's place in the LDS buffer, at the same relative offset;
  * 16-byte lanes are 16-byte aligned on both sides and dword lanes 4-byte
    aligned (global offsets relative to a 16-byte aligned cell block);
  * the pieces cover their LDS places completely and do not overlap;
  * two buffers fit gfx950's 160 KB of LDS per workgroup and the per-wave
    instruction counts fit the 6-bit vmcnt together with one cell's stores.
"""
import os
import subprocess
import tempfile

import pytest

from exwave_amd import build as B

PROGRAM = r'''
#include <cstdio>
#include <vector>
#include "dg_operator_pipe.hpp"
using namespace exw;

static int fails = 0;
#define CHECK(c, ...) do { if (!(c)) { ++fails; std::printf(__VA_ARGS__); std::printf("\n"); } } while (0)

template <int DIM, int K>
void check() {
  using P = PipeShape<DIM, K>;
  if constexpr (!P::ENABLED) return;
  else {
    const DmaPlan &p = P::PLAN;
    const int ob = P::OWN_BYTES, fb = P::FLD_BYTES, cb = P::NCOEF * 8;
    // LDS places of the pieces
    std::vector<int> cover(p.bytes, 0), owner(p.bytes, -1);
    auto place = [&](int id, int lo, int len) {
      CHECK(lo >= 0 && lo + len <= p.bytes, "dim %d k %d: piece %d outside the buffer", DIM, K, id);
      for (int b = lo; b < lo + len && b < p.bytes; ++b) {
        CHECK(owner[b] < 0, "dim %d k %d: pieces %d and %d overlap at %d", DIM, K, owner[b], id, b);
        owner[b] = id;
      }
    };
    place(100, 0, ob);
    place(101, ob, ob);
    for (int j = 0; j < 4 * DIM; ++j) place(j, p.slot[j], fb);
    place(102, p.coef, cb);
    CHECK(2 * p.bytes <= 163840, "dim %d k %d: %d bytes of LDS", DIM, K, 2 * p.bytes);
    CHECK(p.bytes % 16 == 0, "dim %d k %d: buffer not a 16-byte multiple", DIM, K);
    if (P::X4) CHECK(ob % 16 == 0 && cb % 16 == 0, "dim %d k %d: blocks not 16-byte strided", DIM, K);
    for (int i = 0; i < p.n; ++i) {
      const DmaOp &o = p.op[i];
      const int w = o.x4 ? 16 : 4, len = o.lanes * w;
      CHECK(o.lanes >= 1 && o.lanes <= 64, "dim %d k %d op %d: %d lanes", DIM, K, i, o.lanes);
      CHECK(o.goff % w == 0 && o.loff % w == 0, "dim %d k %d op %d: misaligned", DIM, K, i);
      int g0, g1, l0;  // source piece [g0, g1) relative to its block; its LDS place
      if (o.src == DMA_OWN) { g0 = 0; g1 = ob; l0 = 0; }
      else if (o.src == DMA_ACC) { g0 = 0; g1 = ob; l0 = ob; }
      else if (o.src == DMA_COEF) { g0 = 0; g1 = cb; l0 = p.coef; }
      else {
        const int field = (o.fld & 1) ? DIM : (o.fld >> 2);
        g0 = field * fb; g1 = g0 + fb; l0 = p.slot[o.fld];
      }
      CHECK(o.goff >= g0 && o.goff + len <= g1, "dim %d k %d op %d: reads outside its piece", DIM, K, i);
      CHECK(o.loff - l0 == o.goff - g0, "dim %d k %d op %d: lands at the wrong place", DIM, K, i);
      for (int b = o.loff; b < o.loff + len && b < p.bytes; ++b) cover[b] = 1;
    }
    for (int b = 0; b < p.bytes; ++b)
      if (owner[b] >= 0) CHECK(cover[b], "dim %d k %d: byte %d of piece %d not loaded", DIM, K, b, owner[b]);
    CHECK(P::max_count(true) + P::NF <= 63, "dim %d k %d: vmcnt overflow", DIM, K);
    std::printf("dim %d k %d: %d instructions, %d LDS bytes per buffer\n", DIM, K, p.n, p.bytes);
  }
}

template <int DIM, int K>
void all() {
  check<DIM, K>();
  if constexpr (K > 1) all<DIM, K - 1>();
}

int main() {
  all<2, 8>();
  all<3, 8>();
  std::printf("%s\n", fails ? "FAIL" : "ok");
  return fails ? 1 : 0;
}
'''


def test_dma_plans_stay_in_bounds_and_cover():
    with tempfile.TemporaryDirectory() as d:
        srcf = os.path.join(d, "plan.hip")
        exe = os.path.join(d, "plan")
        with open(srcf, "w") as f:
            f.write(PROGRAM)
        r = subprocess.run([B.hipcc(), "-std=c++17", "-O1", f"--offload-arch={B.ARCH}",
                            "-I", B.CSRC, srcf, "-o", exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout[-3000:]
        assert "dim 3 k 4" in r.stdout
