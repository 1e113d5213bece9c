"""Launch table of member groups of different sizes (yelmo_amd/csrc/
ygpu_gtable.h), compiled with g++ and checked on the CPU.  A wrong table
would leave a group's workgroups on several XCDs, on overlapping slots or
waiting for a rank that never arrives, so every invariant the group kernel
relies on is checked:

  every group's workgroups sit on one XCD (b % 8), on consecutive slots
  (b / 8), with ranks 0 .. G-1 once each and the group's size G;
  no slot holds two workgroups; the launch fits the 8 x per_xcd slots;
  every member belongs to exactly one group, of its own size;
  one member per group whenever all groups fit at once.
"""
import os
import random
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = r"""
#include <cstdio>
#include "ygpu_gtable.h"
int main() {
  int per_xcd, n;
  if (std::scanf("%d %d", &per_xcd, &n) != 2) return 2;
  std::vector<int> mg(n);
  for (int k = 0; k < n; ++k) std::scanf("%d", &mg[k]);
  std::vector<int> gtab, gmoff, gmem;
  int nwg = 0;
  std::string err;
  if (!yg::group_table(mg, per_xcd, gtab, gmoff, gmem, nwg, err)) {
    std::printf("ERR %s\n", err.c_str());
    return 0;
  }
  std::printf("%d\n", nwg);
  for (int v : gtab) std::printf("%d ", v);
  std::printf("\n%zu\n", gmoff.size());
  for (int v : gmoff) std::printf("%d ", v);
  std::printf("\n");
  for (int v : gmem) std::printf("%d ", v);
  std::printf("\n");
  return 0;
}
"""


@pytest.fixture(scope="module")
def driver():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "drv.cpp")
    with open(src, "w") as f:
        f.write(DRIVER)
    exe = os.path.join(d, "drv")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "yelmo_amd", "csrc"),
                    src, "-o", exe], check=True)
    return exe


def _table(exe, mgroup, per_xcd):
    inp = f"{per_xcd} {len(mgroup)} " + " ".join(map(str, mgroup))
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    if out.startswith("ERR"):
        return None
    lines = out.split("\n")
    nwg = int(lines[0])
    gtab = [int(v) for v in lines[1].split()]
    gmoff = [int(v) for v in lines[3].split()]
    gmem = [int(v) for v in lines[4].split()] if len(lines) > 4 else []
    return nwg, gtab, gmoff, gmem


def _check(exe, mgroup, per_xcd=64):
    t = _table(exe, mgroup, per_xcd)
    assert t is not None, mgroup
    nwg, gtab, gmoff, gmem = t
    ngrp = len(gmoff) - 1
    assert nwg % 8 == 0 and nwg <= 8 * per_xcd
    assert len(gtab) == 3 * nwg
    wg = {}
    for b in range(nwg):
        g, r, G = gtab[3 * b: 3 * b + 3]
        if g < 0:
            continue
        assert 0 <= g < ngrp and 0 <= r < G
        wg.setdefault(g, []).append((b % 8, b // 8, r, G))
    assert sorted(wg) == list(range(ngrp))
    for g, ws in wg.items():
        G = ws[0][3]
        assert all(w[3] == G for w in ws)
        assert len({w[0] for w in ws}) == 1, "group spans XCDs"
        assert sorted(w[2] for w in ws) == list(range(G))
        slots = sorted(w[1] for w in ws)
        assert slots == list(range(slots[0], slots[0] + G)), "slots not consecutive"
        for m in gmem[gmoff[g]:gmoff[g + 1]]:
            assert mgroup[m] == G
    assert sorted(gmem) == list(range(len(mgroup)))
    if sum(mgroup) <= 8 * per_xcd:
        assert ngrp == len(mgroup)
    return nwg, ngrp


def test_uniform_ensemble_is_the_formula_placement(driver):
    """64 members of G = 4: group g on XCD g % 8, slots 4 (g / 8) ..., as
    the kernel's formula places them without a table."""
    nwg, gtab, gmoff, gmem = _table(driver, [4] * 64, 64)
    _check(driver, [4] * 64)
    assert nwg == 256
    for b in range(nwg):
        slot = b // 8
        assert gtab[3 * b] == (slot // 4) * 8 + b % 8
        assert gtab[3 * b + 1] == slot % 4
    assert gmem == list(range(64))


def test_groups_spread_over_the_xcds(driver):
    """Groups take the XCDs in turn: no XCD holds two more slots than
    another when the sizes are equal."""
    nwg, gtab, _, _ = _table(driver, [4] * 24, 64)
    per = [sum(1 for b in range(nwg) if b % 8 == x and gtab[3 * b] >= 0) for x in range(8)]
    assert max(per) - min(per) <= 4


def test_uneven_diva_batch(driver):
    sides = [96, 128, 192, 256, 384]
    rng = random.Random(0)
    cells = [s * s for s in (rng.choice(sides) for _ in range(48))]

    def g_of(c):
        for g in (16, 8, 4, 2):
            if c // g >= 4096:
                return g
        return 1
    mg = [g_of(c) for c in cells]
    nwg, ngrp = _check(driver, mg)
    assert ngrp == 48


@pytest.mark.parametrize("case", [
    [16] * 40,                    # exceeds 512 slots: groups take members in turn
    [4] * 256,                    # 256 members of G = 4 in two rounds
    [16, 1, 1, 2, 8, 4] * 30,     # every size, more than fit
    [1] * 600,
    [16],
    [2, 16, 4, 8, 1] * 3,
])
def test_tables(driver, case):
    _check(driver, case)


def test_random_mixtures(driver):
    rng = random.Random(1)
    for _ in range(200):
        n = rng.randint(1, 400)
        mg = [rng.choice([1, 2, 4, 8, 16]) for _ in range(n)]
        per_xcd = rng.choice([16, 32, 64])
        _check(driver, mg, per_xcd)


def test_bad_sizes_rejected(driver):
    assert _table(driver, [3, 4], 64) is None
    assert _table(driver, [32], 16) is None
