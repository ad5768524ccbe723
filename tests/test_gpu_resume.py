"""Resume and capacity-matrix files through the host facade on the GPU.

- save_resume writes the reference's resume layout (RootGridIO.cpp:25-290,
  Environment.cpp:379-438), read back here with libfopis_io.so: names,
  shapes, the particle count in "size", generated_counts, phi in the glue
  layout, E of the edge shapes, empty B, the charge map [S][ncmat];
- a run opened with job_type "load" restores particles, phi, the charge map
  and the timestep exactly, and its next step matches the next step of the
  run that saved (no injection, so no random draws: FP64 positions to 1e-12
  of the box, charge map to 1e-10);
- the capacity matrix cache is <cache>/objects/<obj file>.h5 with
  capacity_matrix, grid_spacing and time_spacing (IO.cpp:28-120); a second
  open reads it back bit for bit, and one with another dx rebuilds."""
import json
import os

import numpy as np
import pytest

from tests.cases import CONFIGS, canon, load_config
from tests.test_io import F64, I32, U64, IO

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(CONFIGS), "libfopis_io.so")),
                                 reason="libfopis_io.so not built (no HDF5)")]

F = pytest.importorskip("fopis_amd.lib")
from fopis_amd.host import Host  # noqa: E402


def _cfg(tmp_path, job="new", dx=None, name="in.json"):
    cfg = load_config("c1")
    cfg["Environment"]["nx"] = cfg["Environment"]["ny"] = cfg["Environment"]["nz"] = 17
    cfg["Environment"]["job_type"] = job
    if dx is not None:
        cfg["Environment"]["dx"] = dx
    for v in cfg["Plasma"].values():
        v["particle_per_cell"] = 4
    cfg["Object"] = {"Spacecraft": {"file_name": os.path.join(CONFIGS, "objects", "cube4.obj"),
                                    "is_potential_fixed": False}}
    p = tmp_path / name
    p.write_text(json.dumps(cfg))
    return str(p), cfg


def test_save_load_resume_continues_the_run(tmp_path):
    path, cfg = _cfg(tmp_path)
    a = Host(path, build_cmatrix=True)
    for _ in range(3):
        a.main_loop_es()
    rdir = str(tmp_path / "resume")
    a.save_resume(rdir)
    io = IO()
    h = io.open(os.path.join(rdir, "environment.h5"), 0)
    assert io.read(h, "dx", np.float64, F64) == cfg["Environment"]["dx"]
    assert io.read(h, "dt", np.float64, F64) == cfg["Environment"]["dt"]
    assert int(io.read(h, "timestep", np.int32, I32)) == a.timestep() == 4
    io.L.fopis_io_close(h)
    h = io.open(os.path.join(rdir, "grid_id_0.h5"), 0)
    names = sorted(cfg["Plasma"])  # the facade's species order: ambient species by name
    for s, nm in enumerate(names):
        n = a.ctx.num_particles(s)
        assert io.read(h, f"{nm}/size", np.float64, F64) == n
        assert io.info(h, f"{nm}/x") == (n,)
        gc = io.read(h, f"{nm}/generated_counts", np.uint64, U64)
        assert gc.shape == (6,) and gc[3] == gc[4] == gc[5] == 16 ** 3 * 4  # the initial load's velocities
    assert io.info(h, "Potential/phi") == (19, 19, 19)
    assert io.info(h, "Efield/ex") == (18, 19, 19)
    assert io.info(h, "Efield/ez") == (19, 19, 18)
    assert io.info(h, "Bfield/bx") == (0, 0, 0)
    ncmat = a.object(0)["ncmat"]
    assert io.info(h, "Spacecraft/charge_map") == (len(names), ncmat)
    phi_file = io.read(h, "Potential/phi", np.float64, F64)
    io.L.fopis_io_close(h)
    np.testing.assert_array_equal(phi_file, a.ctx.get_field(F.PHI))

    lpath, _ = _cfg(tmp_path, job="load", name="load.json")
    b = Host(lpath, build_cmatrix=True, resume_dir=rdir)
    assert b.timestep() == a.timestep()
    np.testing.assert_array_equal(b.ctx.get_field(F.PHI), a.ctx.get_field(F.PHI))
    for s in range(len(names)):
        np.testing.assert_array_equal(canon(b.ctx.get_particles(s)), canon(a.ctx.get_particles(s)))
    np.testing.assert_array_equal(b.ctx.object_state(0)["charge_map"], a.ctx.object_state(0)["charge_map"])
    for s in range(len(names)):  # the species' charge densities from the restored particles
        ra = a.ctx.get_field(F.RHO, s + 1)
        np.testing.assert_allclose(b.ctx.get_field(F.RHO, s + 1), ra, rtol=0, atol=1e-12 * np.abs(ra).max())
    a.main_loop_es()
    b.main_loop_es()
    for s in range(len(names)):
        pa, pb = canon(a.ctx.get_particles(s)), canon(b.ctx.get_particles(s))
        assert len(pa) == len(pb)
        assert np.max(np.abs(pa - pb)) <= 1e-12 * 16
    ca, cb = a.ctx.object_state(0)["charge_map"], b.ctx.object_state(0)["charge_map"]
    np.testing.assert_allclose(cb, ca, rtol=0, atol=1e-10 * np.abs(ca).max())
    a.close()
    b.close()


def test_capacity_matrix_cache_in_reference_format(tmp_path):
    path, cfg = _cfg(tmp_path)
    cache = str(tmp_path / "cm")
    a = Host(path, build_cmatrix=True, initialize=False, cmatrix_cache=cache)
    C0 = a.ctx.get_cmatrix(0)
    a.close()
    f = os.path.join(cache, "objects", "cube4.obj.h5")
    assert os.path.exists(f)
    io = IO()
    h = io.open(f, 0)
    n = C0.shape[0]
    assert io.info(h, "capacity_matrix") == (n, n)
    np.testing.assert_array_equal(io.read(h, "capacity_matrix", np.float64, F64), C0)
    assert io.read(h, "grid_spacing", np.float64, F64) == cfg["Environment"]["dx"]
    assert io.read(h, "time_spacing", np.float64, F64) == cfg["Environment"]["dt"]
    io.L.fopis_io_close(h)
    # a second open with build_cmatrix off must take C from the file
    b = Host(path, build_cmatrix=False, initialize=False, cmatrix_cache=cache)
    np.testing.assert_array_equal(b.ctx.get_cmatrix(0), C0)
    b.close()
    # another dx invalidates the file (exact comparison, IO.cpp:91-93): without building, C stays unset
    p2, _ = _cfg(tmp_path, dx=cfg["Environment"]["dx"] * (1 + 1e-12), name="dx2.json")
    c = Host(p2, build_cmatrix=False, initialize=False, cmatrix_cache=cache)
    assert not np.any(c.ctx.get_cmatrix(0))
    c.close()
