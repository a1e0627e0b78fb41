"""Parity of boiling, steam, residues and puffs on the device with the oracle
(particles.comp:98-134, 303-305, 505-517, 633-731, 771-807; grid_prep.comp:163-167, 247;
heat.comp:95-104; air.glsl): pots on a hot floor whose particles start near their boiling
point. Water turns into steam that rises, condenses and escapes as puffs riding the air;
brine with steam off boils away and leaves salt; a fuming oil boils away into puffs. The
same particles change kind, leave and stay on the device as in the oracle (every draw
depends only on ids, the substep count and the particle's own state), the rest agree at the
tolerance of the other whole-substep parity tests, and the ledger counts agree.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu

STEPS = 16


def run(case_kw, steps=STEPS, grid=(64, 48)):
    p = Pair(grid, case=cases.boiling_pot, **case_kw)
    try:
        p.gpu.record(steps, steps / 240)
        p.gpu.wait()
        p.ora.begin_batch(steps / 240, steps)
        for k in range(steps):
            p.ora.substep(1 if k > 0 else 0)
        # every slot, puffs included (downloads leave puffs out: they are no longer part of the tank)
        pa = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, p.gpu.slots())
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pa = pa[pa["kind"] != 0]
        pb = pb[pb["kind"] != 0]
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        led = p.gpu.field("LEDGER", np.uint32, N.LEDGER_WORDS)[16:16 + 3 * N.MAX_MATERIALS]
        lo = p.ora.field("LEDGER", np.uint32, N.LEDGER_WORDS)[16:16 + 3 * N.MAX_MATERIALS]
        return p, pa, pb, led, lo
    except BaseException:
        p.close()
        raise


def check(pa, pb, led, lo):
    assert np.array_equal(pa["id"], pb["id"])      # the same particles left the tank
    assert np.array_equal(pa["kind"], pb["kind"])  # and the same turned into steam, salt, puffs
    assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
    assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
    assert rel_l2(pa["t"], pb["t"]) <= 1e-5
    assert np.array_equal(led, lo)


def test_water_pot_makes_steam_that_rises_condenses_and_escapes():
    p, pa, pb, led, lo = run(dict())
    try:
        assert p.g.caps & N.CAP_STEAM
        low = pb["kind"] & 255
        made, escaped = int(lo[2 * N.MAX_MATERIALS + N.ROW_WATER]), int(lo[N.ROW_STEAM])
        print(f"\nwater pot: steam made {made}, steam now {(low == N.ROW_STEAM + 1).sum()}, "
              f"escaped {escaped}, puffs {(pb['kind'] == N.KIND_PUFF).sum()}", flush=True)
        assert made > 100 and escaped > 0 and (low == N.ROW_STEAM + 1).sum() > 100
        check(pa, pb, led, lo)
    finally:
        p.close()


def test_brine_without_steam_boils_away_and_leaves_salt():
    p, pa, pb, led, lo = run(dict(row=N.ROW_BRINE, steam=0, x_range=(0.9, 1.1)))
    try:
        assert not p.g.caps & N.CAP_STEAM
        salt = int(lo[N.MAX_MATERIALS + N.ROW_SALT])
        print(f"\nbrine: boiled away {int(lo[N.ROW_BRINE])}, salt left {salt}", flush=True)
        assert lo[N.ROW_BRINE] > 100 and salt > 10
        check(pa, pb, led, lo)
    finally:
        p.close()


def test_fuming_oil_boils_away_into_puffs():
    p, pa, pb, led, lo = run(dict(row=N.ROW_OIL, fume=0.5, boil=150.0, latent_vap=300.0e3, x_range=(0.9, 1.1)))
    try:
        puffs = pb["kind"] == N.KIND_PUFF
        print(f"\nfuming oil: boiled away {int(lo[N.ROW_OIL])}, puffs in the air {puffs.sum()}", flush=True)
        assert puffs.sum() > 100
        assert np.all(pb["c"][puffs][:, 0] == np.float32(0.5))
        check(pa, pb, led, lo)
    finally:
        p.close()
