"""Parity of the mouse tools on the device with the oracle (particles.comp:594-631,
air.glsl:24-37): a dam break stepped with each tool under the cursor (grab, erase, push,
pull, the stream nozzle, heat, cool, whirl, hold) loses the same particles and ends within
the tolerance of the other whole-substep parity tests of the oracle stepped with the same
tables; the record's tool reaches the tables (its speed bounds the time step, as
ToolState::speed gives it).
"""
import math

import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu

STEPS = 8
TOOLS = [
    ("grab", 1, (40.0, 10.0), 40.0),
    ("erase", 2, (0.0, 0.0), 0.0),
    ("push", 3, (25.0, 0.0), 40.0),
    ("pull", 3, (-25.0, 0.0), 40.0),
    ("nozzle", 4, (30.0, -12.0), 1000.0),
    ("heat", 5, (0.0, 0.0), 0.0),
    ("cool", 6, (0.0, 0.0), 0.0),
    ("whirl", 7, (-30.0, 0.0), 30.0),
    ("hold", 8, (0.0, 0.0), 600.0),
]


@pytest.mark.parametrize("name,mode,vel,strength", TOOLS, ids=[t[0] for t in TOOLS])
def test_tool_matches(name, mode, vel, strength):
    thermal = mode in (5, 6)
    if thermal:
        p = Pair((64, 48), case=cases.thermal_pool, temp=lambda x, y: np.full(len(x), 40.0))
    else:
        p = Pair((64, 48))
    try:
        W, H = p.grid
        tool = N.ToolState(mode=mode, radius=7.0, strength=strength)
        tool.pos[0], tool.pos[1] = 0.25 * W, 0.2 * H
        tool.vel[0], tool.vel[1] = vel
        p.gpu.record(STEPS, STEPS / 240, tool)
        p.gpu.wait()
        g, mats = p.gpu.grid_params()
        assert g.tool_mode == mode and g.tool_radius == pytest.approx(7.0)
        speed = {1: math.hypot(*vel) + 4.0 * 7.0, 3: math.hypot(*vel), 4: math.hypot(*vel), 7: abs(vel[0]) + 2.0 * 7.0}
        assert g.ext_speed == pytest.approx(speed.get(mode, 0.0), rel=1e-6)
        p.ora.set_gparams(g, mats)
        p.ora.begin_batch(STEPS / 240, STEPS)
        for k in range(STEPS):
            p.ora.substep(1 if k > 0 else 0)
        pa = p.gpu.download_particles()
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = pb[pb["kind"] != 0]
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        assert np.array_equal(pa["id"], pb["id"])
        if mode == 2:
            assert len(pb) < len(p.parts) - 50  # the brush took liquid away
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
        if thermal:
            assert rel_l2(pa["t"], pb["t"]) <= 1e-6
            # the tool's energy in the ledger (entry 4 heat, 5 cool), the same fixed-point sums
            a = p.gpu.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16]
            b = p.ora.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16]
            k = 4 if mode == 5 else 5
            qa = int(a[2 * k]) | int(a[2 * k + 1]) << 32
            qb = int(b[2 * k]) | int(b[2 * k + 1]) << 32
            assert qb > 0 and abs(qa - qb) <= 1e-5 * qb
            assert p.gpu.download_ledger()[k] > 0.0
    finally:
        p.close()
