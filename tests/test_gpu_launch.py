"""Launch shapes of the step kernel on the GPU: the statistics steps time their own
shapes (uniform z segmentations of 8-row tiles), and the choice changes no result."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_stats_launch_timing_keeps_results_bitwise(mcg, storage):
    """Jet 130x70x98 with statistics on every step: the first statistics steps
    time the candidate shapes (each step another one) and keep the fastest; the
    fields and the radial sums after 60 steps equal those of a run with 3 fixed
    segments bit for bit, and the timing has chosen."""
    c = mcg.perturbed_jet_case(130, 70, 98, jet_radius=8.0)
    c.NSTATS_START_STEP = 0
    c.NSTATS_SAMPLE = 1
    out = []
    for seg in (0, 3):
        print(f"jet 130x70x98 {storage}, statistics every step, segments {seg or 'timed'}", flush=True)
        s = mcg.Solver(c, storage=storage)
        s.set_launch_segments(seg)
        s.init_jet()
        s.step(60)
        f = s.fields()
        r = s.radial()
        out.append(({k: f[k] for k in ("rhor", "rhob", "ux", "uy", "uz")}, {k: r[k] for k in r}, s.launch_info()))
        s.close()
    print(f"launch shapes: {[o[2] for o in out]}", flush=True)
    timed, fixed = out
    assert timed[2]["stats"]["tuned"] and not fixed[2]["stats"]["tuned"]
    assert not timed[2]["tuned"]  # no plain step ran
    for k in timed[0]:
        assert np.array_equal(timed[0][k], fixed[0][k]), k
    for k in timed[1]:
        assert np.array_equal(np.asarray(timed[1][k]), np.asarray(fixed[1][k])), k
