"""Writes tests/golden/karma_n16.npz: the Karma reference program
(karma/karma3d.f90, compiled from its source into oracle/_ref/ref_karma3d by
oracle/ref/Makefile) run on a 16^3 grid with noise -- its Tecplot output at
steps 20 and 40 parsed into arrays, the files' header records and names, and
its console transcript -- and a second run of the same namelist with
tip_stop = 0.3, which the tip passes at the first output step (the
program's early exit).

    make -C oracle/ref && python tests/golden/make_karma_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
EXE = os.path.join(ROOT, "oracle", "_ref", "ref_karma3d")

# karma.nml for the run (the other entries keep the program's defaults)
NML = dict(N=16, nsteps=40, nout=20, Fnoise=0.05, tip_stop=2.0, R0=6.0)


def read_tecplot(path, n):
    rows = np.loadtxt(path, skiprows=4)
    assert rows.shape == (n ** 3, 6), rows.shape
    # POINT order: x fastest, then y, then z
    return rows[:, 3].reshape(n, n, n), rows[:, 5].reshape(n, n, n)


def _run(nml):
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "karma.nml"), "w") as f:
            f.write("&karma\n" + "".join(f" {k} = {v}\n" for k, v in nml.items()) + "/\n")
        os.makedirs(os.path.join(d, "output"))
        env = dict(os.environ, OMP_NUM_THREADS="1", OMP_TARGET_OFFLOAD="DISABLED")
        r = subprocess.run([EXE], cwd=d, check=True, env=env, stdout=subprocess.PIPE, text=True)
        names = sorted(os.listdir(os.path.join(d, "output")))
        fields, headers = {}, []
        for fn in names:
            it = int(fn[len("karma_"):-len(".dat")])
            path = os.path.join(d, "output", fn)
            with open(path) as fh:
                headers.append([next(fh).rstrip("\n") for _ in range(4)])
            fields[it] = read_tecplot(path, nml["N"])
        return r.stdout.splitlines(), names, headers, fields


def main():
    out = {}
    lines, names, headers, fields = _run(NML)
    for it in (20, 40):
        out[f"phase_{it}"], out[f"conc_{it}"] = fields[it]
    out["stdout_lines"] = np.array(lines)
    out["filenames"] = np.array(names)
    out["headers"] = np.array(headers)
    # the early exit: same namelist, tip_stop = 0.3 (tip at i = 8 > 4.8 at it = 20)
    stop = dict(NML, tip_stop=0.3)
    lines, names, _, _ = _run(stop)
    out["stop_stdout_lines"] = np.array(lines)
    out["stop_filenames"] = np.array(names)
    out["stop_nml"] = np.array(list(stop.items()), dtype=object).astype(str)
    np.savez_compressed(os.path.join(HERE, "karma_n16.npz"),
                        nml=np.array(list(NML.items()), dtype=object).astype(str), **out)
    print("wrote karma_n16.npz", file=sys.stderr)


if __name__ == "__main__":
    main()
