"""Host side of the similarity source order (beampower_amd/csrc/bp_order.h).

The header is compiled here with g++ into a small shared object, so these
tests need no GPU.  Bar: the order is a permutation whose tiles are whole
(but the last) and list their sources in ascending index; the result on
any number of host threads is identical to the one-thread result; it
narrows the tiles of a shuffled grid.
"""
import ctypes as ct
import os
import shutil
import subprocess

import numpy as np
import pytest

from beampower_amd import synthetic

HEADER_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                          "beampower_amd", "csrc")
SHIM = r'''
#include "bp_order.h"
extern "C" void order(const int *tau, long K, long D, long tile, unsigned threads, int *out) {
  auto o = threads ? bp::similarity_order(tau, K, D, tile, threads)
                   : bp::similarity_order(tau, K, D, tile);
  for (long i = 0; i < K; ++i) out[i] = o[i];
}
'''


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    d = tmp_path_factory.mktemp("order")
    src, so = d / "shim.cpp", d / "shim.so"
    src.write_text(SHIM)
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", HEADER_DIR,
                    str(src), "-o", str(so)], check=True)
    lib = ct.CDLL(str(so))
    lib.order.argtypes = [ct.c_void_p, ct.c_long, ct.c_long, ct.c_long, ct.c_uint, ct.c_void_p]
    lib.order.restype = None

    def run(tau, tile=48, threads=1):
        tau = np.ascontiguousarray(tau, dtype=np.int32)
        K = tau.shape[0]
        D = int(np.prod(tau.shape[1:])) if tau.ndim > 1 else 0
        out = np.empty(K, np.int32)
        lib.order(tau.ctypes.data, K, D, tile, threads, out.ctypes.data)
        return out
    return run


def tile_spreads(tau, order, tile):
    """Widest delay range over a tile's sources, per tile."""
    t = tau.reshape(tau.shape[0], -1)[order]
    return np.array([int((t[i:i + tile].max(0) - t[i:i + tile].min(0)).max())
                     for i in range(0, len(t), tile)])


def shuffled(seed=3, grid=(24, 20, 10), S=16):
    rng = np.random.default_rng(seed)
    _, _, tau = synthetic.make_network(seed, S, 90.0, grid)
    tau = np.ascontiguousarray(tau[rng.permutation(tau.shape[0])])
    # some identical rows, so that the split keys tie
    tau[rng.choice(tau.shape[0], 200, replace=False)] = tau[0]
    return tau


@pytest.mark.parametrize("tile", [48, 64])
def test_order_is_a_permutation_of_sorted_whole_tiles(shim, tile):
    tau = shuffled(grid=(23, 19, 10))  # 4,370 sources: the last tile is short
    o = shim(tau, tile)
    K = tau.shape[0]
    assert np.array_equal(np.sort(o), np.arange(K))
    for i in range(0, K, tile):
        part = o[i:i + tile]
        # ascending index within each tile: leaves start at multiples of the
        # tile (a tile holding parts of two leaves would in general not be sorted)
        assert np.all(np.diff(part) > 0), i


@pytest.mark.parametrize("threads", [2, 3, 8, 16])
def test_threads_give_the_one_thread_order(shim, threads):
    tau = shuffled(5, grid=(30, 30, 12))
    one = shim(tau, 48, 1)
    assert np.array_equal(shim(tau, 48, threads), one)
    assert np.array_equal(shim(tau, 48, 0), one)  # the library's own thread count


def test_order_narrows_a_shuffled_grid(shim):
    tau = shuffled(7)
    K = tau.shape[0]
    nat = tile_spreads(tau, np.arange(K), 48)
    sim = tile_spreads(tau, shim(tau, 48), 48)
    assert sim.mean() < 0.5 * nat.mean()
    assert np.median(sim) < np.median(nat)


def test_order_of_small_or_empty_inputs_is_the_identity(shim):
    tau = shuffled()
    assert np.array_equal(shim(tau[:48], 48), np.arange(48))
    assert np.array_equal(shim(tau[:1], 48), np.arange(1))
    assert np.array_equal(shim(np.zeros((100, 0), np.int32), 48), np.arange(100))
    # identical rows: ties go by index, so the order is the identity
    same = np.repeat(tau[:1], 500, axis=0)
    assert np.array_equal(shim(same, 48), np.arange(500))
