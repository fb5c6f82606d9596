"""Host-side API checks that need no GPU: the RAW loader's file-name parsing and
value conversions (rawFile.cpp:31-128), the C++ header compiling on its own,
the offline driver's argument handling (viewer.cpp:70-161, 204-224), and the
host-side deal of the persistent batched launch's work items."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def test_rawr_parses_name_and_converts(tmp_path):
    from mdwtracker_amd.renderer import RawR
    v = (np.arange(2 * 3 * 4) % 7).astype(np.int8).reshape(4, 3, 2)
    p = tmp_path / "blob_2x3x4_int8.raw"
    v.tofile(p)
    r = RawR(p)
    assert r.getDims() == (2, 3, 4) and r.getBytesPerVoxel() == 1
    np.testing.assert_array_equal(r.pushable(), v.astype(np.float32) * 255 - 128)  # rawFile.cpp:102
    u = np.arange(8, dtype=np.uint8).reshape(2, 2, 2)
    q = tmp_path / "x_2x2x2_uint8.raw"
    u.tofile(q)
    ru = RawR(q)
    assert ru.pushable().dtype == np.uint8  # u8 goes to the GPU as stored (x255 inside)
    ru.reshapeBounds()
    assert ru.bounds[1].tolist() == [2.0, 2.0, 2.0]
    with pytest.raises(ValueError):
        RawR(tmp_path / "nodims_uint8.raw")


def test_cpp_header_compiles_standalone(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "mdwt_renderer.hpp"\nint main() { return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


DEAL_CHECK = r"""
#include <algorithm>
#include <cstdio>
#include <random>
#include "batch_deal.hpp"
int main() {
  std::mt19937_64 g(7);
  for (int trial = 0; trial < 200; ++trial) {
    const size_t n = trial < 10 ? size_t(trial) : 1 + g() % 3000;
    const size_t G = 1 + g() % 300;
    std::vector<unsigned long long> t(2 * n);
    for (size_t i = 0; i < n; ++i) {
      const unsigned long long a = g() % 1000000;
      const int kind = trial % 4;
      const unsigned long long c = kind == 0 ? g() % 16 : kind == 1 ? g() % 300000
                                 : kind == 2 ? g() : (g() % 3 == 0 ? 0 : g() % 5000);
      t[2 * i] = a;
      t[2 * i + 1] = a + c;
      if (kind == 3 && g() % 7 == 0) std::swap(t[2 * i], t[2 * i + 1]);
    }
    std::vector<size_t> pos(n);
    for (size_t i = 0; i < n; ++i) pos[i] = i;
    std::stable_sort(pos.begin(), pos.end(), [&](size_t x, size_t y) {
      return mdwt::item_cost(t.data(), x) > mdwt::item_cost(t.data(), y);
    });
    const std::vector<uint32_t> order = mdwt::order_by_decreasing_cost(t.data(), n);
    const std::vector<uint32_t> src = mdwt::deal_sources(t.data(), n, G);
    std::vector<int> seen(n, 0);
    for (size_t idx = 0; idx < n; ++idx) {
      if (order[idx] != pos[idx]) { printf("order differs: trial %d at %zu\n", trial, idx); return 1; }
      const size_t r = idx / G, w = idx % G, cnt = std::min(G, n - r * G);
      if (src[idx] != pos[r * G + ((r & 1) ? cnt - 1 - w : w)]) { printf("deal differs\n"); return 1; }
      ++seen[src[idx]];
    }
    for (size_t i = 0; i < n; ++i)
      if (seen[i] != 1) { printf("not a permutation\n"); return 1; }
    // balance: any two workgroups' sums differ by at most the largest cost
    std::vector<unsigned long long> sum(G, 0ull);
    unsigned long long cmax = 0;
    for (size_t idx = 0; idx < n; ++idx) {
      const unsigned long long c = mdwt::item_cost(t.data(), src[idx]);
      sum[idx % G] += c;
      cmax = std::max(cmax, c);
    }
    const auto mm = std::minmax_element(sum.begin(), sum.end());
    if (trial % 4 != 2 && *mm.second - *mm.first > cmax) { printf("unbalanced\n"); return 1; }
  }
  printf("ok\n");
  return 0;
}
"""


def test_batch_deal_orders_like_a_stable_sort(tmp_path):
    """Launch form 2's re-deal: the radix order by decreasing cost equals a
    stable comparison sort's (ties, zero and out-of-order stamps, 64-bit
    costs), the dealt list is a permutation in alternating rounds of G, and
    any two workgroups' cost sums differ by at most the largest cost (except
    for full 64-bit costs, whose sums overflow)."""
    src = tmp_path / "deal.cpp"
    src.write_text(DEAL_CHECK)
    exe = tmp_path / "deal"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror",
                        f"-I{ROOT / 'mdwtracker_amd' / 'csrc'}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr


def test_offline_cli_argument_errors():
    exe = ROOT / "mdwtracker_amd" / "mdwt_offline"
    if not exe.exists():
        pytest.skip("mdwt_offline not built")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "No data file given" in r.stderr
    r = subprocess.run([str(exe), "-fr", "a_4x4x4_uint8.raw", "-mc", "0"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "positive" in r.stderr
    r = subprocess.run([str(exe), "-fr", "a_4x4x4_uint8.raw", "-l", "1", "0", "0"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "-l takes 5 to 7 values" in r.stderr


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parent.parent
    p = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_bench_batch_views_are_seeded_and_bounded():
    import sys
    from pathlib import Path

    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import bench

    a, b = bench.batch_views(64, 7), bench.batch_views(64, 7)
    assert a == b and len(a) == 64
    for v in a:
        assert 256 <= v["width"] <= 1024 and 256 <= v["height"] <= 1024 and 1 <= v["spp"] <= 16
        d = [t - f for t, f in zip(v["at"], v["from"])]
        norm = sum(x * x for x in d) ** 0.5
        # the up vector is never along the view direction
        assert abs(sum(x * u for x, u in zip(d, v["up"])) / norm) < 0.99
