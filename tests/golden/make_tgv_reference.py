"""Writes tgv_filtered_reference_64_32.json: the filtered kinetic energy and spectra of the
plain FP64 Taylor-Green run (64^3 elements, N = 4, Mach 0.1, Re 1600, molecular viscosity,
17,304 LSRK54 steps to t* = 8) on the 32^3 coarse brick at t* = 1, ..., 8, read by the
configuration-4 a-posteriori test.  Needs a GPU (about 6 minutes on one MI355X).

    python tests/golden/make_tgv_reference.py [OUT.json]
"""
import os
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "climatemachine.jl_amd"))

from cmdg import tgv_closure  # noqa: E402

if __name__ == "__main__":
    out = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "tgv_filtered_reference_64_32.json")
    with tempfile.TemporaryDirectory() as tmp:
        tgv_closure.main(["--stage", "eval", "--fine", "64", "--coarse", "32", "--tend", "8",
                          "--eval-times", "1,2,3,4,5,6,7,8", "--models", "", "--save-reference", out, "--out", tmp])
