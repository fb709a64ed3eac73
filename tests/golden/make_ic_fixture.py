"""SHA-256 digests of the compiled reference's initial states for the cases
of tests/golden/ic_cases.py (oracle/_ref/libref.so: set_initial_conditions_3d
and set_initial_conditions_3d_from_deck on the reference's own solver).

    python tests/golden/make_ic_fixture.py     (seconds; writes ic_problems.npz)
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import ic_cases  # noqa: E402
import oracle_lib as OL  # noqa: E402


def reference_state(ref, problem, shape, box, gamma, G, deck_path):
    s = ref.ref_create(*shape, *box, G)
    ref.ref_set_gamma(s, gamma)
    if deck_path:
        ref.ref_ic_deck(s, deck_path.encode())
    else:
        ref.ref_ic_problem(s, problem)
    U = np.zeros(tuple(shape) + (5,))
    ref.ref_get_state(s, U.ctypes.data_as(OL._dp))
    ref.ref_destroy(s)
    return U


def main():
    ref = OL.ref_lib()
    assert ref is not None, "build oracle/_ref first (python -c 'import __graft_entry__ as g; g.build()')"
    out = {}
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory() as d:
        ic_cases.prepare(d)
        os.chdir(d)
        try:
            for name, problem, shape, box, gamma, G, deck in ic_cases.CASES:
                path = None
                if deck:
                    path = os.path.join(d, name + ".deck")
                    with open(path, "w") as f:
                        f.write(deck)
                U = reference_state(ref, problem, shape, box, gamma, G, path)
                out[name] = np.frombuffer(hashlib.sha256(U.tobytes()).digest(), dtype=np.uint8)
        finally:
            os.chdir(cwd)
    np.savez(os.path.join(HERE, "ic_problems.npz"), **out)
    print("wrote ic_problems.npz:", ", ".join(out))


if __name__ == "__main__":
    main()
