"""The division-free log-odds quotient of the plain score kernel equals the
IEEE quotient on the operand ranges the kernel admits (Markstein's
correction; the C check is compiled here with gcc, whose fmaf is the
correctly rounded fused multiply-add the GPU's v_fma_f32 is)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_quotient_identity_on_admitted_ranges(tmp_path):
    exe = str(tmp_path / "quotient_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-o", exe,
                    os.path.join(HERE, "quotient_check.c"), "-lm"], check=True)
    r = subprocess.run([exe, "20000000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("0 mismatched"), r.stdout
