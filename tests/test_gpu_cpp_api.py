"""GPU runs of the C++ programs written against the drop-in headers
(include/mallocMC): the reference's example01 / example03 usage and its
verify_heap harness (tests/verify_heap.cpp of the reference), built by
__graft_entry__.build() into tests/cpp/_build/."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BUILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_build")


def _run(prog, *args, timeout=120):
    path = os.path.join(BUILD, prog)
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: run __graft_entry__.build()")
    print(f"running {prog} {' '.join(args)}", flush=True)
    r = subprocess.run([path, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def test_example_arrays():
    r = _run("example_arrays")
    assert r.returncode == 0 and "example_arrays OK" in r.stdout


def test_example_slots():
    r = _run("example_slots")
    assert r.returncode == 0 and "example_slots OK" in r.stdout


@pytest.mark.parametrize("heap_mb,blocks,threads", [(1024, 64, 128), (256, 256, 256)])
def test_verify_heap(heap_mb, blocks, threads):
    r = _run("verify_heap", str(heap_mb), str(blocks), str(threads))
    assert r.returncode == 0 and "verify_heap OK" in r.stdout


def test_user_written_policies():
    """A distribution policy and an OOM policy written against the
    reference's policy concept (device_allocator.hpp:55-107) plug into the
    Allocator: requests rounded to 64 B take one 64-byte slot each, every
    null reaches the user's handleOOM on the per-lane and the wave path."""
    r = _run("custom_policies")
    assert r.returncode == 0 and "custom_policies OK" in r.stdout


def test_udiv24_exhaustive():
    """The float-reciprocal division of pdiv equals n / d for every n <= 65536
    and every d in 1..65536 (every pagesize up to 64 KiB, every divisor the
    heap uses), and on 2^28 sampled pairs below 2^24 wherever the quotient
    is below 2^22, and clamped to 32 always."""
    r = _run("udiv24_exhaustive")
    assert r.returncode == 0 and "udiv24 OK" in r.stdout


def test_thread_malloc_drop_in():
    """1M threads malloc and free through the drop-in handle: full
    wavefronts take the wave-cooperative path, wavefronts with a lane missing
    the per-lane path; both keep the blocks intact and restore the slots."""
    r = _run("thread_malloc", timeout=150)
    assert r.returncode == 0 and "thread_malloc OK" in r.stdout
