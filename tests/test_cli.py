"""grid.out / dft.out / wstack.out argument handling (no GPU needed for these paths)."""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "radio-gridders_amd")


def run(exe, *args):
    return subprocess.run([os.path.join(PKG, exe), *args], capture_output=True, text=True,
                          timeout=60)


def test_help_and_missing_arguments():
    r = run("grid.out", "--help")
    assert r.returncode == 0 and "-wkernel" in r.stdout
    r = run("grid.out", "-theta=0.1")
    assert r.returncode == 1 and "No Visibility File" in r.stdout
    r = run("grid.out", "-vis=a", "-wkernel=b", "-theta=0.1", "-lambda=100")
    assert r.returncode == 1 and "Missing W-Towers Parameters" in r.stdout
    r = run("dft.out", "-vis=a", "-theta=0.1", "-lambda=100")
    assert r.returncode == 1 and "blocks" in r.stdout
    r = run("dft.out", "-HELP")
    assert r.returncode == 0 and "-threadblock" in r.stdout
    r = run("wstack.out", "-help")
    assert r.returncode == 0 and "-sepkern_uv" in r.stdout
    r = run("wstack.out", "-theta=0.1", "-lambda=20480")
    assert r.returncode == 1 and "sepkern_uv" in r.stdout
    r = run("wstack.out", "-theta=0.1", "-lambda=20480", "-sepkern_uv=a", "-sepkern_w=b",
            "-sepkern_lm=c", "-sepkern_n=d", "-mode=0")
    assert r.returncode == 2
