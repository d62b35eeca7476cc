"""The EnsNVnBath command line (src/application/EnsNVnBath.cpp:112-140):
option names, short flags, defaults and units; config.cfg; the dat/ layout;
on the GPU, a run equals the batch API plus the CCE reduction."""
import os

import numpy as np
import pytest

import oops_amd as O
from oops_amd import ensnvnbath as E

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_defaults_match_the_reference_application():
    para, _ = E.parse([])
    assert para["input"] == "C13Bath/RoyCoord.xyz" and para["logfile"] == "EnsembleCCE.conf"
    assert para["position"] == "0.0 0.0 0.0" and para["state0"] == 0 and para["state1"] == 1
    assert para["cce"] == 3 and para["cutoff"] == 6.0 and para["polarization"] == "0.0 0.0 0.0"
    assert para["dephasing_rate"] == 0.0 and para["dephasing_axis"] == "1.0 1.0 1.0"
    assert para["nTime"] == 101 and para["start"] == 0.0 and para["finish"] == 0.002
    assert para["magnetic_field"] == "0.1 0.1 0.1" and para["pulse"] == "CPMG" and para["pulse_num"] == 1
    assert para["output"] == "EnsNVnBath"


def test_short_flags_and_output_name(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    argv = ["-c", "4", "-d", "8", "-n", "41", "-f", "1e-3", "-B", "0.05 0.05 0.05", "-m", "2", "-a", "0", "-b", "2",
            "-r", "10", "-P", "1 2 3"]
    para, _ = E.parse(argv)
    assert (para["cce"], para["cutoff"], para["nTime"], para["finish"], para["pulse_num"]) == (4, 8.0, 41, 1e-3, 2)
    assert para["state1"] == 2 and para["dephasing_rate"] == 10.0
    assert np.array_equal(E._vec(para["magnetic_field"]), [0.05, 0.05, 0.05])
    assert np.array_equal(E._vec(para["position"]), [1.0, 2.0, 3.0])
    # the reference names the result after the concatenated arguments
    assert para["output"] == "EnsNVnBath" + "".join(argv)


def test_config_file_fills_options_not_on_the_command_line(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    (tmp_path / "config.cfg").write_text("# comment\ncce = 2\ncutoff=7.5\nnTime = 11\n")
    para, _ = E.parse(["-c", "4"])
    assert para["cce"] == 4 and para["cutoff"] == 7.5 and para["nTime"] == 11


def test_initialize_creates_the_dat_layout(tmp_path, monkeypatch):
    monkeypatch.setenv("CCE_PROJ_PATH", str(tmp_path))
    assert E.main(["-I"]) == 0
    for d in ("log", "input", "output", "config", "debug"):
        assert (tmp_path / "dat" / d).is_dir()


def _project(tmp_path, monkeypatch):
    monkeypatch.setenv("CCE_PROJ_PATH", str(tmp_path))
    E.main(["-I"])
    X = O.bath_diamond(3.567, 16.0, 0.011, seed=2)
    path = str(tmp_path / "dat" / "input" / "bath.xyz")
    O.write_xyz(path, X, ["13C"] * len(X))
    return O.read_xyz(path)[0]    # the file holds the coordinates to its text precision


@pytest.mark.gpu
@pytest.mark.parametrize("npulse", [1, 4])
def test_run_equals_batch_and_reduction(ctx, tmp_path, monkeypatch, npulse):
    X = _project(tmp_path, monkeypatch)
    argv = ["-i", "bath.xyz", "-c", "3", "-d", "8", "-n", "21", "-f", "1e-3", "-B", "0.05 0.05 0.05",
            "-m", str(npulse), "-o", "res"]
    para, _ = E.parse(argv)
    out, res = E.run(para, ctx=ctx, verbose=False)
    z = np.load(out, allow_pickle=False)
    C = O.Clusters(X, 8.0, 3)
    par = O.cce_params(B=[0.05, 0.05, 0.05])
    t = np.linspace(0.0, 1e-3, 21)
    b = C.echo_batch(ctx, par, npulse, t / 2.0 if npulse == 1 else t, m=30, tol=1e-12, precision="fp64")
    b.run()
    Ls, st = b.results()
    want = C.reduce(np.stack(Ls).reshape(-1), 21)["final"]
    print(f"EnsNVnBath N={npulse}: clusters {C.counts()}, L(t_end) {z['final_result'][-1]}", flush=True)
    assert np.abs(z["final_result"] - want).max() <= 1e-12
    assert np.array_equal(z["time_list"][:, 0], t)
