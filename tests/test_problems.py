"""Problem generators: the five workloads' parameters and invariants."""
import math

import numpy as np
import pytest

from snap_amd import problems


def test_config0_tables():
    p = problems.config0()
    p.validate()
    assert (p.nx, p.ny, p.nz, p.ng, p.nang, p.nmom) == (16, 16, 16, 4, 10, 1)
    assert np.allclose(p.sigt[0], 1.0 + 0.01 * np.arange(4))
    assert np.allclose(p.siga, 0.5 * p.sigt) and np.allclose(p.sigs, p.sigt - p.siga)
    assert np.all(p.qi == 1.0) and (p.epsi, p.iitm, p.oitm) == (1e-4, 5, 100)
    assert p.units_per_sweep == 16 ** 3 * 80 * 4


def test_scattering_split_conserves_and_downscatters():
    sigs = np.array([0.5, 0.6, 0.7, 0.8, 0.9])
    m = problems.scattering_split(sigs, 3)
    # every source group scatters exactly sigs[g] in total (column sums)
    assert np.allclose(m[:, :, 0].sum(axis=0), sigs, rtol=1e-15)
    # half stays in group, none goes up, the last group keeps everything
    assert np.allclose(np.diag(m[:, :, 0])[:-1], 0.5 * sigs[:-1])
    assert m[-1, -1, 0] == sigs[-1]
    assert np.all(np.triu(m[:, :, 0], 1) == 0.0)
    assert np.allclose(m[:, :, 2], 0.01 * m[:, :, 0])


def test_config1_layout():
    p = problems.config1(n=16, ng=6, nang=8, nmom=4)
    p.validate()
    assert p.cmom == 16 and p.nmat == 2
    cube = p.mat == 1
    assert cube.sum() == 8 ** 3 and cube[4:12, 4:12, 4:12].all()
    assert np.all(p.qi[:, cube] == 1.0) and np.all(p.qi[:, ~cube] == 0.0)
    assert np.allclose(p.siga[1], 2.0 * p.siga[0]) and np.allclose(p.sigs[1], p.sigs[0])
    assert np.allclose(p.sigt[1], p.siga[1] + p.sigs[1])
    assert np.allclose(p.slgg[:, :, :, 1], 0.1 * p.slgg[:, :, :, 0])


def test_config3_lognormal_field():
    p = problems.config3(n=24, ng=3, nang=4)
    p.validate()
    siga = p.sigt_cell[0] - 0.5
    logs = np.log(siga)
    assert abs(logs.mean() - math.log(0.5)) < 0.02
    assert abs(logs.std() - 0.3) < 0.05
    assert np.all(p.sigt_cell == p.sigt_cell[0])       # the same field in every group
    assert p.qi[:, :8, :8, :8].min() == 1.0 and p.qi.sum() == 3 * 8 ** 3
    # deterministic from the seed
    assert np.array_equal(problems.config3(n=24, ng=3, nang=4).sigt_cell, p.sigt_cell)


def test_config4_batch():
    ps = problems.config4(count=12)
    assert len(ps) == 12
    for p in ps:
        p.validate()
        assert {p.nx, p.ny, p.nz} <= {24, 32, 40, 48, 56, 64}
        assert (p.ng, p.nang, p.nmom) == (16, 24, 4)
        assert 0 < (p.mat == 1).sum() < p.ncells
    sizes = {(p.nx, p.ny, p.nz) for p in ps}
    assert len(sizes) > 1


def test_rng_is_stable():
    u = problems.splitmix64_uniform(1, 4)
    assert np.all((u > 0) & (u < 1))
    assert np.array_equal(u, problems.splitmix64_uniform(1, 4))
    assert not np.array_equal(u, problems.splitmix64_uniform(2, 4))
    z = problems.standard_normal(3, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


def test_validate_rejects_bad_tables():
    p = problems.config0(n=4, ng=2, nang=2)
    p.mat[0, 0, 0] = 3
    with pytest.raises(ValueError):
        p.validate()
    p = problems.config0(n=4, ng=2, nang=2)
    p.epsi = 0.5
    with pytest.raises(ValueError):
        p.validate()


def test_bench_algorithmic_bytes_per_unit():
    """The roofline's algorithmic bytes per cell x direction x group of the
    reference data flow: dinv streamed once per octant (8 B, 4 B in FP32),
    plus the source moments, t_xs and the flux-moment read-modify-write
    (3*8*cmom + 8 B per cell, group and octant) shared by the octant's
    angles -- 11.2, 18.89 and 9.33 B for the 16^3 P0 10-angle, the 64^3 P3
    36-angle and the 48^3 P0 24-angle workloads."""
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location(
        "bench", os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    assert math.isclose(bench.bytes_per_unit(10, 1, 8), 11.2)
    assert math.isclose(bench.bytes_per_unit(36, 16, 8), 18.0 + 8.0 / 9.0)
    assert math.isclose(bench.bytes_per_unit(24, 1, 8), 8.0 + 32.0 / 24.0)
    assert math.isclose(bench.bytes_per_unit(36, 16, 4), 14.0 + 8.0 / 9.0)
    # units of one sweep of every octant and group
    p = problems.config1()
    assert p.units_per_sweep == 64 ** 3 * 8 * 36 * 32
