"""Geometry on the GPU against the host builders (which are pinned to the
reference's golden vectors): masks and counts bit-exact."""
import numpy as np
import pytest
import torch

import prtlb
from prtlb import geometry as G
from conftest import golden

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,radius,phi,seed", [
    ((128, 128), 6.0, 0.6, 0), ((97, 61), 3.5, 0.55, 3),
    ((48, 40, 32), (4.0, 8.0), 0.35, 1), ((33, 17, 21), (2.0, 5.0), 0.45, 7),
    ((64, 64, 64), (3.0, 8.0), 0.25, 11)])
def test_device_pack_equals_host(shape, radius, phi, seed):
    print("case", shape, radius, phi, seed, flush=True)
    for connect in (False, True):
        host = G._pack(shape, radius, phi, seed, connect, None)
        dev = G._pack(shape, radius, phi, seed, connect, "cuda")
        assert np.array_equal(host, dev), connect


def test_device_connectivity_and_coating_match_reference():
    d = golden("geometry")
    mt = d["meta"]
    for g, kept, rem in ((d["g"], d["kept"], mt["removed"]), (d["g2d"], d["kept2d"], mt["removed2d"])):
        t = torch.from_numpy(np.ascontiguousarray(g)).cuda()
        t, n = G.keep_inlet_connected_device(t)
        assert np.array_equal(t.cpu().numpy(), kept) and n == rem
    for layers, code, key, cnt in ((1, 3, "coat1", "n1"), (2, 4, "coat2", "n2")):
        t = torch.from_numpy(np.ascontiguousarray(d["kept"])).cuda()
        t, n = G.coat_with_biofilm_device(t, code, layers)
        assert np.array_equal(t.cpu().numpy(), d[key]) and n == mt[cnt]


def test_device_connectivity_random_large():
    rng = np.random.default_rng(5)
    g = np.where(rng.random((96, 80, 72)) < 0.32, 0, 2).astype(np.uint8)
    host, nh = prtlb.keep_inlet_connected(g)
    t, nd = G.keep_inlet_connected_device(torch.from_numpy(g).cuda())
    assert nh == nd and np.array_equal(host, t.cpu().numpy())
