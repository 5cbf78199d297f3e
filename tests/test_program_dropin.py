"""The whole reference program with its solve on the GPU.

tests/fortran/_build/bsam_gpu is the reference program -- BSAMDriver, the
regrid every step (error tagging, Berger-Rigoutsos, InitFields,
TransferValues), the time-step loop and the frame writer, all compiled
unmodified from the reference sources -- with its MultigridIterations call
sent to the engine at link time. Run on the same input deck as the reference
program itself (oracle/_ref/bsam), every frame file out/mNNNNN.dat must be
byte-identical: the hierarchy every step (patch lists, mglobal, mx) and q
with its ghost layer, as printed.
"""

import os

import pytest

import bsam_program as bp

pytestmark = pytest.mark.gpu


def first_difference(a: bytes, b: bytes) -> str:
    la, lb = a.decode().splitlines(), b.decode().splitlines()
    for n, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {n + 1}: gpu {x!r} reference {y!r}"
    return f"lengths {len(la)} vs {len(lb)} lines"


def variant(name: str, steps: int):
    deck = bp.shipped_deck(steps)
    g = deck["griddata"]
    if name == "neumann_l3":
        g["mthbc"] = {i: 1 for i in range(1, 7)}
        g["maxlevel"] = 3
    if name == "3d_l1":
        # the 3D problem (QInit3D, SetErrFlagsUser3D, f25.12 frames) on a
        # 32 x 16 x 32 root, periodic in x, user BC elsewhere, one level
        g.update(ndims=3, maxlevel=1, minlevel=-3, mx={1: 32, 2: 16, 3: 32},
                 mglobal={(1, 1): 1, (1, 2): 32, (2, 1): 1, (2, 2): 16, (3, 1): 1, (3, 2): 32})
        deck["rundata"]["dt"] = 0.01
    return deck


@pytest.mark.parametrize("name,steps", [("shipped", 6), ("neumann_l3", 3), ("3d_l1", 3)])
def test_program_frames_identical(tmp_path, name, steps):
    for prog in (bp.REF_PROGRAM, bp.GPU_PROGRAM):
        if not os.path.exists(prog):
            pytest.skip(f"{prog} not built (needs the reference build)")
    deck = variant(name, steps)
    ref, _, t_ref = bp.run(bp.REF_PROGRAM, deck, str(tmp_path / "ref"))
    gpu, out, t_gpu = bp.run(bp.GPU_PROGRAM, deck, str(tmp_path / "gpu"))
    print(f"{name}: {steps} steps, reference program {t_ref:.2f} s, "
          f"with the GPU solve {t_gpu:.2f} s")
    assert sorted(ref) == list(range(steps + 1)), sorted(ref)
    assert sorted(gpu) == sorted(ref), (sorted(gpu), out[-2000:])
    for n in sorted(ref):
        assert gpu[n] == ref[n], f"frame {n}: {first_difference(gpu[n], ref[n])}"
