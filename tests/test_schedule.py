"""Banded launch schedules of the step kernel (multicomponent_color_gradient_amd/csrc/schedule.hpp),
built on the host: every tile-plane is run exactly once, the XCDs get equal shares, an XCD runs
its units band by band with the long bands first, and the estimated span obeys the
list-scheduling bounds.  The header is plain C++, compiled here with g++; no GPU is used."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
XCDS = 8


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("sched") / "schedule_dump")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", os.path.join(HERE, "schedule_dump.cpp"), "-o", exe],
                   check=True)

    def run(ntiles, iz, cus, fill):
        out = subprocess.run([exe, str(ntiles), str(iz), str(cus), str(fill)], check=True, capture_output=True,
                             text=True).stdout.split("\n")
        cands = []
        for line in out:
            f = line.split()
            if not f:
                continue
            if f[0] == "cand":
                cands.append({"bands": int(f[1]), "n": int(f[2]), "span": float(f[3]), "deal": int(f[4]), "units": []})
            elif f[0] == "th":
                cands[-1]["th"] = [int(v) for v in f[1:]]
            elif f[0] == "u":
                cands[-1]["units"].append(tuple(int(v) for v in f[1:]))
        return cands

    return run


# (tiles, planes, CUs per XCD): the 256x512x256 jet with 12-row and 8-row
# tiles, the 128x256x128 jet, a tall 130x130x5002 box, and small grids
SHAPES = [(172, 254, 32, 1.9), (256, 254, 32, 1.5), (44, 126, 32, 1.9), (22, 5000, 32, 1.9), (6, 96, 32, 1.9),
          (3, 20, 4, 1.5)]


@pytest.mark.parametrize("ntiles,iz,cus,fill", SHAPES)
def test_banded_schedules_cover_every_tile_plane_once(dump, ntiles, iz, cus, fill):
    cands = dump(ntiles, iz, cus, fill)
    # up to three band rules dealt in contiguous shares, then the best one paired and mirrored
    assert 3 <= len(cands) <= 5
    assert [c["deal"] for c in cands] == [0] * (len(cands) - 2) + [1, 2]
    assert cands[-1]["th"] == cands[-2]["th"] == cands[0]["th"]
    for c in cands:
        u = c["units"]
        assert len(u) == c["n"] and len(u) % XCDS == 0
        hits = np.zeros((ntiles, iz + 1), dtype=np.int32)
        for tile, zb, ze in u:
            if ze < zb:
                assert (tile, zb, ze) == (0, 1, 0)  # padding
                continue
            assert 0 <= tile < ntiles and 1 <= zb <= ze <= iz
            hits[tile, zb:ze + 1] += 1
        assert np.all(hits[:, 1:] == 1)
        assert np.all(hits[:, 0] == 0)


@pytest.mark.parametrize("ntiles,iz,cus,fill", SHAPES)
def test_banded_xcd_shares_and_band_order(dump, ntiles, iz, cus, fill):
    W = ntiles * iz
    for c in dump(ntiles, iz, cus, fill):
        u = c["units"]
        th = c["th"]
        assert len(th) == c["bands"] >= 2 and sum(th) == iz and min(th) >= 1
        # thickness decreases, except that the last band takes the rest
        assert all(a >= b for a, b in zip(th[:-2], th[1:-1]))
        edges = np.concatenate([[0], np.cumsum(th)])
        for k in range(XCDS):
            mine = u[k::XCDS]
            real = [v for v in mine if v[2] >= v[1]]
            # padding only after the XCD's last unit
            assert mine[:len(real)] == real
            # contiguous: an equal share of the tile-major tile-planes; paired and
            # mirrored: two of 16 equal chunks, k and k+8, or k and 15-k
            if c["deal"] == 0:
                want = list(range(W * k // XCDS, W * (k + 1) // XCDS))
            else:
                other = k + XCDS if c["deal"] == 1 else 2 * XCDS - 1 - k
                want = [i for ch in (k, other) for i in range(W * ch // (2 * XCDS), W * (ch + 1) // (2 * XCDS))]
            idx = sorted(t * iz + z - 1 for t, zb, ze in real for z in range(zb, ze + 1))
            assert idx == sorted(want)
            # band by band, tiles in order within a band; a unit lies within one band
            band = [int(np.searchsorted(edges, ze, side="left")) - 1 for _, _, ze in real]
            for (t, zb, ze), j in zip(real, band):
                assert edges[j] < zb <= ze <= edges[j + 1]
            # (a tile may hold two units of one band where an XCD's two chunks meet in it)
            order = [(j, t, zb) for (t, zb, _), j in zip(real, band)]
            assert order == sorted(order)


@pytest.mark.parametrize("ntiles,iz,cus,fill", SHAPES)
def test_banded_span_within_list_scheduling_bounds(dump, ntiles, iz, cus, fill):
    W = ntiles * iz
    for c in dump(ntiles, iz, cus, fill):
        real = [v for v in c["units"] if v[2] >= v[1]]
        per_xcd = [[v for v in c["units"][k::XCDS] if v[2] >= v[1]] for k in range(XCDS)]
        work = [sum(ze - zb + 1 + fill for _, zb, ze in p) for p in per_xcd]
        longest = max(ze - zb + 1 + fill for _, zb, ze in real)
        # Graham's bound for list scheduling on cus machines, per XCD
        lo = max(w / cus for w in work)
        hi = max(w / cus + (1 - 1 / cus) * max([ze - zb + 1 + fill for _, zb, ze in p] or [0]) for w, p in
                 zip(work, per_xcd))
        assert lo - 1e-6 <= c["span"] <= hi + 1e-6, (c["span"], lo, hi)
        assert c["span"] >= W / XCDS / cus
        assert c["span"] <= W / XCDS / cus + longest + fill * len(real) / XCDS


def test_banded_candidates_beat_uniform_at_the_jet_size(dump):
    """At 256x512x256 with 12-row tiles (172 tiles, 254 planes, 32 CUs per XCD) the
    best banded schedule's estimated span is below that of every uniform
    segmentation (the best, 7 segments of 37 planes, leaves 30% of the CUs idle
    through its last round)."""
    fill = 1.9
    best = dump(172, 254, 32, fill)[0]["span"]
    for zs in range(1, 65):
        zlen = -(-254 // zs)
        nzs = -(-254 // zlen)
        n = 172 * nzs
        # units of one length dealt to XCDs in contiguous ranges: the busiest XCD runs ceil(n/8) of them
        units = -(-n // XCDS)
        rounds = -(-units // 32)
        span = rounds * (zlen + fill)
        assert best < span, (zs, span, best)
