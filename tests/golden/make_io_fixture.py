"""Fixture for the run loop and the files at the boundary: the compiled
reference (oracle/_ref/libref.so, built from its own sources) runs a small
self-gravitating deck through run_simulation_3d with CSV and VTK output,
then a second solver restarts from that run's output_000001.csv through
load_from_csv_restart and run_simulation_3d_restart (its potential starts
cold, as the reference's does).  Recorded: SHA-256 and size of every output
file, and the lines the loops print.

    python tests/golden/make_io_fixture.py      (writes tests/golden/io_run.json)

Each reference run happens in a child process (the loops print through C
stdio) whose working directory holds the deck, the profile file and output/.
Shared with tests/test_gpu_io.py: DECK, RESTART_TIME, run_lines().
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "io_run.json")

N = 16
# the octant explosion of the reference's example deck, small, with outputs
DECK = f"""problem = STELLAR_OCTANT_EXPLOSION_3D
nx = {N}
ny = {N}
nz = {N}
xmin = 0.0
xmax = 1.5
ymin = 0.0
ymax = 1.5
zmin = 0.0
zmax = 1.5
t_end = 0.03
output_dt = 0.01
cfl = 0.4
gamma = 1.666667
G_const = 1.0
bc_xmin = REFLECTIVE
bc_xmax = OUTFLOW
bc_ymin = REFLECTIVE
bc_ymax = OUTFLOW
bc_zmin = REFLECTIVE
bc_zmax = OUTFLOW
stellar_profile = stellar_profile.txt
explosion_energy = 0.1
explosion_radius = 0.2
explosion_type = thermal
"""
# keys only this build reads: the reference's moment order, so that the
# device's self-gravity steps are the reference's bits
DEVICE_KEYS = "moment_order = sequential\n"
RESTART_FILE = "output_000001.csv"
RESTART_TIME = 0.01
RESTART_COUNT = 1

# the lines of the run loops (hydro_solver_3d.c:2241-2330) that are compared
LINE_PREFIXES = ("Starting 3D simulation", "Restarting 3D simulation", "Starting output count",
                 "Grid:", "Step ", "Simulation completed")


def run_lines(text):
    return [ln for ln in text.splitlines() if ln.startswith(LINE_PREFIXES)]


def digests(directory):
    out = {}
    for name in sorted(os.listdir(directory)):
        with open(os.path.join(directory, name), "rb") as f:
            data = f.read()
        out[name] = {"sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data)}
    return out


CHILD = r"""
import ctypes as C, sys
lib = C.CDLL(sys.argv[1])
lib.input_deck_create.restype = C.c_void_p
lib.input_deck_create.argtypes = [C.c_char_p]
lib.solver3d_create_from_deck.restype = C.c_void_p
lib.solver3d_create_from_deck.argtypes = [C.c_void_p]
lib.set_initial_conditions_3d_from_deck.argtypes = [C.c_void_p, C.c_void_p]
lib.run_simulation_3d.argtypes = [C.c_void_p, C.c_double, C.c_double]
lib.run_simulation_3d_restart.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int]
lib.load_from_csv_restart.argtypes = [C.c_void_p, C.c_char_p]
lib.load_from_csv_restart.restype = C.c_int
deck = lib.input_deck_create(b"run.deck")
s = lib.solver3d_create_from_deck(deck)
t_end, out_dt = float(sys.argv[3]), float(sys.argv[4])
if sys.argv[2] == "run":
    lib.set_initial_conditions_3d_from_deck(s, deck)
    lib.run_simulation_3d(s, t_end, out_dt)
else:
    assert lib.load_from_csv_restart(s, sys.argv[5].encode()) == 0
    lib.run_simulation_3d_restart(s, float(sys.argv[6]), t_end, out_dt, int(sys.argv[7]))
C.CDLL(None).fflush(None)
"""


def reference_run(ref_so, workdir, mode, restart_from=None):
    sys.path.insert(0, HERE)
    import ic_cases
    ic_cases.prepare(workdir)
    with open(os.path.join(workdir, "run.deck"), "w") as f:
        f.write(DECK)
    os.makedirs(os.path.join(workdir, "output"), exist_ok=True)
    args = [sys.executable, "-c", CHILD, ref_so, mode, "0.03", "0.01"]
    if mode == "restart":
        args += [restart_from, repr(RESTART_TIME), str(RESTART_COUNT)]
    r = subprocess.run(args, cwd=workdir, capture_output=True, text=True, check=True)
    return run_lines(r.stdout)


def main():
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libref.so")
    if not os.path.exists(ref_so):
        sys.exit("build the reference first: make -C oracle")
    sys.path.insert(0, ROOT)
    with tempfile.TemporaryDirectory() as tmp:
        run1 = os.path.join(tmp, "run")
        os.makedirs(run1)
        lines1 = reference_run(ref_so, run1, "run")
        files1 = digests(os.path.join(run1, "output"))
        run2 = os.path.join(tmp, "restart")
        os.makedirs(run2)
        src = os.path.join(run1, "output", RESTART_FILE)
        lines2 = reference_run(ref_so, run2, "restart", src)
        files2 = digests(os.path.join(run2, "output"))
    fixture = {"deck": DECK, "run": {"lines": lines1, "files": files1},
               "restart": {"from": RESTART_FILE, "t_start": RESTART_TIME,
                           "output_count": RESTART_COUNT, "lines": lines2, "files": files2}}
    with open(OUT, "w") as f:
        json.dump(fixture, f, indent=1)
    print(f"wrote {OUT}: {len(files1)} + {len(files2)} files, "
          f"{len(lines1)} + {len(lines2)} lines")


if __name__ == "__main__":
    main()
