"""The shared-reciprocal division of nv_divide.h equals IEEE float division (the kernels rely on it to
reproduce the reference's a/b bit for bit). Built and run on the CPU with the seed reciprocal perturbed by up
to 2 ulp, which covers the hardware reciprocal's 1-ulp error. The cases include quotients within t 2^-49 of a
float midpoint (which a reciprocal good to 2^-46 gets wrong) and exact ties of the subnormal grid (which the
product alone gets wrong without the IEEE fallback)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_shared_reciprocal_division_is_ieee(tmp_path):
    exe = tmp_path / "divcheck"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), os.path.join(HERE, "csrc", "div_shared_reciprocal.c"),
                    "-lm"], check=True)
    r = subprocess.run([str(exe), "20000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0" in r.stdout
