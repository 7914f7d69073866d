"""Pins of the HEALPix NESTED implementation (frt_healpix.h) that stands in for
the un-vendored HEALPix C++ 3.40 library on both the GPU and the CPU checker.

Checked against what Gorski et al. 2005 fixes: the 12 base-pixel centres, the
ring latitudes, the nested hierarchy (children 4p..4p+3 inside the parent),
equal pixel areas, and pix2vec/vec2pix being inverse.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as P


def unit_vectors(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_base_pixel_centres_closed_form():
    # faces 0-3: z = 2/3, phi = pi/4 + k pi/2; faces 4-7: z = 0, phi = k pi/2;
    # faces 8-11: z = -2/3, phi = pi/4 + k pi/2
    for f in range(12):
        v = P.pix2vec(0, f)
        k = f % 4
        if f < 4:
            z, phi = 2.0 / 3.0, math.pi / 4 + k * math.pi / 2
        elif f < 8:
            z, phi = 0.0, k * math.pi / 2
        else:
            z, phi = -2.0 / 3.0, math.pi / 4 + k * math.pi / 2
        st = math.sqrt(1 - z * z)
        assert np.allclose(v, [st * math.cos(phi), st * math.sin(phi), z], atol=1e-15), f


@pytest.mark.parametrize("order", [1, 2, 3])
def test_ring_latitudes(order):
    # ring i (1..4N-1): z = 1 - i^2/(3N^2) (i < N), 4/3 - 2i/(3N) (N <= i <= 3N),
    # mirrored in the south; ring i holds 4i (polar) or 4N (equatorial) pixels
    N = 1 << order
    zs = np.array([P.pix2vec(order, p)[2] for p in range(12 * N * N)])
    expect = []
    for i in range(1, 4 * N):
        if i < N:
            z, cnt = 1 - i * i / (3 * N * N), 4 * i
        elif i <= 3 * N:
            z, cnt = 4 / 3 - 2 * i / (3 * N), 4 * N
        else:
            j = 4 * N - i
            z, cnt = -(1 - j * j / (3 * N * N)), 4 * j
        expect.append((z, cnt))
    for z, cnt in expect:
        assert np.sum(np.abs(zs - z) < 1e-12) == cnt


@pytest.mark.parametrize("order", [0, 1, 2, 3, 4, 5])
def test_roundtrip_all_pixels(order):
    npix = 12 << (2 * order)
    for p in range(npix):
        v = P.pix2vec(order, p)
        assert abs(np.linalg.norm(v) - 1.0) < 1e-15
        assert P.vec2pix(order, v) == p


@pytest.mark.parametrize("order", [8, 10, 12])
def test_roundtrip_sampled_high_order(order):
    rng = np.random.default_rng(order)
    npix = 12 << (2 * order)
    for p in rng.integers(0, npix, 3000):
        assert P.vec2pix(order, P.pix2vec(order, int(p))) == int(p)


def test_nesting_consistency():
    # vec2pix at order l, shifted by 2 bits, is vec2pix at order l-1
    for v in unit_vectors(4000, 3):
        pix = [P.vec2pix(l, v) for l in range(13)]
        for l in range(1, 13):
            assert pix[l] >> 2 == pix[l - 1]


def test_vec2pix_ignores_length():
    for v in unit_vectors(500, 4):
        for l in (0, 3, 7):
            assert P.vec2pix(l, v) == P.vec2pix(l, 3.7 * v)


def test_equal_area():
    # uniform directions land in each of the 192 order-2 pixels equally often
    v = unit_vectors(200000, 5)
    counts = np.bincount([P.vec2pix(2, x) for x in v], minlength=192)
    expect = len(v) / 192
    chi2 = float(((counts - expect) ** 2 / expect).sum())
    # 191 degrees of freedom: mean 191, sd ~19.5; a wrong area would give >> 300
    assert chi2 < 300, chi2


def test_children_inside_parent():
    # the centre of child 4p+c lies within the parent pixel
    for order in range(0, 6):
        for p in range(0, 12 << (2 * order), 7):
            for c in range(4):
                v = P.pix2vec(order + 1, 4 * p + c)
                assert P.vec2pix(order, v) == p
