"""The page claim / page reset protocol of the wave path on CPU threads
(tests/pte_protocol_model.c, compiled with gcc here): over every interleaving
of a freer and two claimers, a claim that reserves with one 64-bit add on the
paired (chunk size, count) word never holds a chunk of a page whose chunk
size changes under it; the earlier split reservation (chunk-size CAS, then a
32-bit count add) does, so the search is able to find the race; 8 threads
claiming and freeing one page show no violation either."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def _build(tmp_path):
    exe = tmp_path / "pte_model"
    subprocess.run(["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pthread",
                    os.path.join(HERE, "pte_protocol_model.c"), "-o", str(exe)], check=True)
    return str(exe)


def test_paired_reservation_survives_every_interleaving(tmp_path):
    r = subprocess.run([_build(tmp_path), "model"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert "model paired: 0 violations" in r.stdout
    split = [ln for ln in r.stdout.splitlines() if ln.startswith("model split:")]
    assert split and int(split[0].split()[2]) > 0, "the split reservation's race must be found"
    assert r.returncode == 0


def test_paired_reservation_on_threads(tmp_path):
    r = subprocess.run([_build(tmp_path), "threads"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert "threads paired: 0 violations" in r.stdout and r.returncode == 0
