"""EnsNVnBath on one MI355X: the ensemble CCE of an NV centre in a nuclear
spin bath with the reference application's command line
(src/application/EnsNVnBath.cpp:112-140): the same option names, short
flags, defaults and units, the same dat/ folder layout and config.cfg file,
and a result file with the reference's variables.

    python -m oops_amd.ensnvnbath -i C13Bath/RoyCoord.xyz -c 4 -d 8 -n 401 -f 5e-3 -B "0.05 0.05 0.05" -p CPMG -m 1

Options (name, short flag, default, unit):
  --input, -i          C13Bath/RoyCoord.xyz   bath file under $CCE_PROJ_PATH/dat/input/ (or ./dat/input/)
  --output, -o         "EnsNVnBath" + the command-line arguments concatenated
                                              result file name under .../dat/output/ (".npz" is appended)
  --logfile, -l        EnsembleCCE.conf       logging configuration (accepted, not used)
  --position, -P       "0.0 0.0 0.0"          NV position, Angstrom
  --state0, -a         0                      first NV state index (eigenstates of the NV Hamiltonian, ascending)
  --state1, -b         1                      second NV state index
  --cce, -c            3                      CCE order
  --cutoff, -d         6.0                    pair cut-off distance of the connection graph, Angstrom
  --polarization, -z   "0.0 0.0 0.0"          bath polarization vector
  --dephasing_rate, -r 0.0                    bath dephasing rate, Hz (times 2 pi in the Liouvillian)
  --dephasing_axis, -x "1.0 1.0 1.0"          bath dephasing axis (normalised)
  --range, -g, --number, -N, --seed, -D, --isotope   accepted, not used (as in the reference)
  --nTime, -n          101                    number of time points
  --start, -s          0.0                    start time, s
  --finish, -f         0.002                  finish time, s
  --magnetic_field, -B "0.1 0.1 0.1"          magnetic field, Tesla
  --pulse, -p          CPMG                   pulse sequence name (CPMG only, as Pulse_Timing)
  --pulse_num, -m      1                      number of pulses N (1: Hahn echo)
  --initialize, -I                            create the dat/ folder layout and exit
  --gpu_precision      fp64                   fp64 (the reference's accuracy) or mixed (not in the reference)

Options not on the command line are read from ./config.cfg ("name = value"
lines) when it exists; the command line wins.  The time list is
linspace(start, finish, nTime) as in the reference (ensemble_cce.cpp:30);
the evolution follows the reference's piecewise propagator, whose state k is
k steps of dt = (finish - start) / (nTime - 1) from rho0
(QuantumEvolutionAlgorithm.cpp:60-83), i.e. the coherence at k dt.

Differences: the result is an .npz with the variables of CCE::export_mat_file
(cce.cpp:190-230) instead of a MATLAB file; the bath must be 13C (the device
kernels take spin-1/2 baths); one process and one GPU instead of MPI ranks.
"""
import argparse
import os
import sys
import time

import numpy as np

# (long name, short flag, default, type); the reference's po::options_description
_OPTIONS = [
    ("input", "i", "C13Bath/RoyCoord.xyz", str),
    ("output", "o", None, str),
    ("logfile", "l", "EnsembleCCE.conf", str),
    ("position", "P", "0.0 0.0 0.0", str),
    ("state0", "a", 0, int),
    ("state1", "b", 1, int),
    ("cce", "c", 3, int),
    ("cutoff", "d", 6.0, float),
    ("polarization", "z", "0.0 0.0 0.0", str),
    ("dephasing_rate", "r", 0.0, float),
    ("dephasing_axis", "x", "1.0 1.0 1.0", str),
    ("range", "g", 1000.0, float),
    ("number", "N", 100, int),
    ("seed", "D", 1, int),
    ("isotope", None, "E", str),
    ("nTime", "n", 101, int),
    ("start", "s", 0.0, float),
    ("finish", "f", 0.002, float),
    ("magnetic_field", "B", "0.1 0.1 0.1", str),
    ("pulse", "p", "CPMG", str),
    ("pulse_num", "m", 1, int),
    ("gpu_precision", None, "fp64", str),
]


def _vec(text):
    """Armadillo vec(string): whitespace-separated numbers."""
    return np.array([float(v) for v in str(text).split()], float)


def _paths():
    root = os.environ.get("CCE_PROJ_PATH") or os.getcwd()
    dat = os.path.join(root, "dat")
    return {k: os.path.join(dat, k) for k in ("log", "input", "output", "config", "debug")}


def parse(argv):
    ap = argparse.ArgumentParser(prog="EnsNVnBath", add_help=False)
    ap.add_argument("-h", "--help", action="store_true")
    ap.add_argument("-I", "--initialize", action="store_true")
    for name, short, _, typ in _OPTIONS:
        flags = ["--" + name] + (["-" + short] if short else [])
        ap.add_argument(*flags, dest=name, type=typ, default=None)
    a = ap.parse_args(argv)
    para = {}
    given = {name for name, *_ in _OPTIONS if getattr(a, name) is not None}
    for name, _, default, typ in _OPTIONS:
        para[name] = getattr(a, name) if name in given else default
    if os.path.exists("config.cfg"):
        types = {name: typ for name, _, _, typ in _OPTIONS}
        for line in open("config.cfg"):
            line = line.split("#", 1)[0].strip()
            if "=" not in line:
                continue
            key, val = (s.strip() for s in line.split("=", 1))
            if key in types and key not in given:
                para[key] = types[key](val)
    if para["output"] is None:
        para["output"] = "EnsNVnBath" + "".join(argv)
    para["help"], para["initialize"] = a.help, a.initialize
    return para, ap


def run(para, ctx=None, verbose=True):
    import oops_amd as O
    paths = _paths()
    if para["pulse"] != "CPMG":
        raise ValueError("Pulse type not supported: " + para["pulse"])
    nt, t0, t1, npulse = para["nTime"], para["start"], para["finish"], para["pulse_num"]
    if nt < 2 or npulse < 1:
        raise ValueError("nTime must be >= 2 and pulse_num >= 1")
    X, iso = O.read_xyz(os.path.join(paths["input"], para["input"]))
    if set(iso) - {"13C"}:
        raise ValueError("the device path takes 13C baths; the file holds " + ", ".join(sorted(set(iso))))
    params = O.cce_params(B=_vec(para["magnetic_field"]), nv_pos=_vec(para["position"]), state0=para["state0"],
                          state1=para["state1"], dephasing_rate=para["dephasing_rate"],
                          dephasing_axis=_vec(para["dephasing_axis"]), pol=_vec(para["polarization"]))
    ctx = ctx or O.Context(0)
    t_start = time.perf_counter()
    C = O.Clusters(X, para["cutoff"], para["cce"])
    dt = (t1 - t0) / (nt - 1)
    tau = dt * np.arange(nt)                     # the reference's k dt from rho0
    prec = para["gpu_precision"]
    if prec not in ("fp64", "mixed"):
        raise ValueError("gpu_precision must be fp64 or mixed")
    b = C.echo_batch(ctx, params, npulse, tau / 2.0 if npulse == 1 else tau, m=30,
                     tol=1e-12 if prec == "fp64" else 1e-3, precision=prec)
    b.run()
    Ls, st = b.results()
    failed = int(((st[:, 8] & 0xFF) != 0).sum())
    time_list = np.linspace(t0, t1, nt)
    L = np.stack(Ls)
    out = os.path.join(paths["output"], para["output"] + ".npz")
    os.makedirs(paths["output"], exist_ok=True)
    C.export_npz(out, L, time_list)
    res = C.reduce(L.reshape(-1), nt)
    if verbose:
        print(f"EnsNVnBath: {len(X)} bath spins, clusters per order {C.counts()}, {failed} failed, "
              f"{time.perf_counter() - t_start:.2f} s; L(t_end) per order {np.round(res['final'][-1], 6).tolist()}; "
              f"result: {out}", flush=True)
    return out, res


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    para, ap = parse(argv)
    if para["help"]:
        print(__doc__)
        return 0
    if para["initialize"]:
        for p in _paths().values():
            os.makedirs(p, exist_ok=True)
        print("dat folder initialized: " + os.path.dirname(_paths()["input"]))
        return 0
    run(para)
    return 0


if __name__ == "__main__":
    sys.exit(main())
