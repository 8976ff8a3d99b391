"""RADMC-3D file interop (pdspy_amd.radmc3d): the formats pdspy writes and
reads (pdspy/radmc3d/write.py, read.pyx:433-478) and the round trip to the
arrays the GPU path takes."""
import numpy as np
import pytest

from pdspy_amd import configs, radmc3d, rt


def _model(nphi=1):
    m = configs.disk_c1(nr=12, nt=10, nphi=nphi)
    g = m.grid
    nr, nt, np_ = len(g.w1) - 1, len(g.w2) - 1, len(g.w3) - 1
    ir, it, ip = np.meshgrid(np.arange(nr), np.arange(nt), np.arange(np_), indexing="ij")
    # distinct integer values per cell pin the axis order; %f keeps them exact
    g.temperature = [10.0 + ir + 100.0 * it + 10000.0 * ip for _ in g.density]
    return m


def test_dust_temperature_round_trip_and_layout(tmp_path):
    m = _model(nphi=3)
    g = m.grid
    T = g.temperature[0]
    path = str(tmp_path / "dust_temperature.dat")
    radmc3d.write_dust_temperature(path, [T, 2 * T])
    lines = open(path).read().split("\n")
    # header and r-fastest order (write.py:266-289)
    assert lines[:3] == ["1", "%d" % T.size, "2"]
    assert lines[3] == "%f" % T[0, 0, 0] and lines[4] == "%f" % T[1, 0, 0]
    assert lines[3 + T.shape[0]] == "%f" % T[0, 1, 0]
    flat = radmc3d.read_dust_temperature(path)
    assert len(flat) == 2 and flat[0].size == T.size
    back = radmc3d.read_temperature_into(m, path)
    assert np.array_equal(back[0], T) and np.array_equal(back[1], 2 * T)
    assert back[0].shape == T.shape


def test_binary_dust_temperature(tmp_path):
    T = np.arange(24, dtype=float).reshape(4, 3, 2) + 0.25
    flat = np.ascontiguousarray(T.transpose(2, 1, 0)).ravel()
    path = str(tmp_path / "dust_temperature.bdat")
    with open(path, "wb") as f:
        np.array([1, 8, flat.size, 2], dtype="<i8").tofile(f)
        np.concatenate([flat, flat * 3]).astype("<f8").tofile(f)
    got = radmc3d.read_dust_temperature(path, binary=True)
    assert np.array_equal(got[0], flat) and np.array_equal(got[1], 3 * flat)
    assert np.array_equal(radmc3d.to_grid_layout(got[0], 4, 3, 2), T)


def test_write_model_files_follow_pdspy_formats(tmp_path):
    m = _model()
    radmc3d.write_model(m, str(tmp_path), nphot_therm=1000000, iseed=3,
                        modified_random_walk=1, scattering_mode_max=0, setthreads=4)
    ctl = open(tmp_path / "radmc3d.inp").read().split("\n")
    assert "nphot_therm = 1000000" in ctl and "iseed = 3" in ctl
    assert "modified_random_walk = 1" in ctl and "setthreads = 4" in ctl
    amr = open(tmp_path / "amr_grid.inp").read().split("\n")
    assert amr[:6] == ["1", "0", "100", "0", "1  1  0", "12  10  1"]
    assert amr[6] == "%12.9e" % (m.grid.w1[0] * 1.495978707e13)
    opac = open(tmp_path / "dustopac.inp").read().split("\n")
    assert opac[:2] == ["2", "1"] and opac[3:6] == ["1", "0", "0"]
    kap = open(tmp_path / "dustkappa_0.inp").read().split("\n")
    assert kap[0] == "2" and len(kap[2].split()) == 3
    st = open(tmp_path / "stars.inp").read().split("\n")
    assert st[0] == "2" and st[1] == "1  %d" % len(m.grid.lam)
    assert st[-2] == "%f" % -m.grid.stars[0].temperature
    # the imaging temperature (rho-weighted species mean) for every species
    T = radmc3d.read_dust_temperature(str(tmp_path / "dust_temperature.dat"))
    assert np.allclose(T[0], rt.imaging_temperature(m.grid), rtol=1e-6)


def test_model_directory_round_trip_to_device_arrays(tmp_path):
    m = configs.disk_envelope(nr=16, nt=12)
    radmc3d.write_model(m, str(tmp_path))
    got = radmc3d.read_model_arrays(str(tmp_path))
    ref = rt.model_arrays(m.grid)
    assert np.allclose(got["rw"], ref["rw"], rtol=1e-9, atol=0)
    assert np.allclose(got["tw"], ref["tw"], rtol=1e-9, atol=1e-12)
    assert np.allclose(got["pw"], ref["pw"], rtol=1e-9, atol=1e-12)
    assert got["rho"].shape == ref["rho"].shape
    assert np.allclose(got["rho"], ref["rho"], rtol=2e-6, atol=0)
    assert np.allclose(got["lam_um"], ref["lam_um"], rtol=1e-5)
    assert np.allclose(got["kabs"], ref["kabs"], rtol=1e-4)
    assert np.allclose(got["ksca"], ref["ksca"], rtol=1e-4)
    assert got["gsca"] is None
    assert np.allclose(got["stars"], ref["stars"], rtol=1e-6)


def test_unsupported_inputs_raise(tmp_path):
    p = tmp_path / "amr_grid.inp"
    p.write_text("1\n0\n0\n0\n1 1 1\n1 1 1\n0\n1\n0\n1\n0\n1\n")
    with pytest.raises(ValueError, match="spherical"):
        radmc3d.read_amr_grid(str(p))
    s = tmp_path / "stars.inp"
    s.write_text("2\n1 2\n1e11 1e33 0 0 0\n1.0\n2.0\n5.0\n6.0\n")
    with pytest.raises(ValueError, match="blackbody"):
        radmc3d._read_stars(str(s))
