"""GPU tests of the device-resident serving loop (include/mallocmc_serve.h).

Every completed request's checksum is read back from its paged buffer on the
heap and must equal the closed form; the same requests run alone must give
the same checksums as in the mix (exact: integer payloads); admission keeps
reserved bytes within the budget and the heap returns to its initial slot
count once every request completed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MiB = 1024 ** 2


def _alloc(mmc, heap=256 * MiB, pagesize=16384, regionsize=16):
    cfg = mmc.ScatterConfig(pagesize=pagesize, accessblocksize=64 * MiB, regionsize=regionsize, resetfreedpages=True,
                            flags=mmc.FLAG_FIX_PAGESIZE_LEAK)
    return mmc.ScatterAllocator(cfg, heap)


def _run_all(loop, trace, chunk=512, max_steps=200000):
    """Steps until every request completed (admission may make them wait)."""
    loop.load(trace)
    done = 0
    while done < max_steps:
        loop.run(chunk, graph=True, time_steps=False)
        done += chunk
        st = loop.stats()
        if st["completed"] == len(trace.arrival):
            return st
    return loop.stats()


# the first two have kernels built for their configuration, the third runs the
# kernels that read every constant from the heap view
@pytest.mark.parametrize("pagesize,regionsize", [(16384, 16), (65536, 32), (32768, 32)])
def test_serve_checksums_and_restore(mmc, cuda, pagesize, regionsize):
    from mallocmc_amd import serve
    a = _alloc(mmc, pagesize=pagesize, regionsize=regionsize)
    s0 = a.getAvailableSlots(4096)
    loop = serve.ServeLoop(a, max_slots=4096, max_tokens=2048)
    trace = serve.make_trace(seed=3, rate_per_s=2000, step_s=0.01, duration_s=1.0, mu=5.0, sigma=0.8,
                             max_tokens=2048)
    print(f"serve trace: {len(trace.arrival)} requests", flush=True)
    st = _run_all(loop, trace)
    assert st["completed"] == len(trace.arrival) and st["nulls"] == 0, st
    assert st["block_mallocs"] == st["block_frees"] == int(np.sum((trace.length + 15) // 16))
    cs = loop.checksums()
    for i in range(len(trace.arrival)):
        assert int(cs[i]) == serve.expected_checksum(int(trace.ids[i]), int(trace.length[i])), i
    cuda.cuda.synchronize()
    assert a.getAvailableSlots(4096) == s0
    loop.close()
    a.close()


def test_serve_alone_equals_mix_under_admission_pressure(mmc, cuda):
    """A budget of 8 MiB forces requests to wait for bytes to free up; the
    checksums of requests run alone equal those in the mix."""
    from mallocmc_amd import serve
    a = _alloc(mmc)
    loop = serve.ServeLoop(a, max_slots=1024, max_tokens=1024, budget_bytes=8 * MiB)
    trace = serve.make_trace(seed=4, rate_per_s=4000, step_s=0.01, duration_s=0.5, mu=5.5, sigma=0.6,
                             max_tokens=1024)
    st = _run_all(loop, trace)
    assert st["completed"] == len(trace.arrival) and st["nulls"] == 0, st
    assert st["queue_wait_steps"] > 0, "the budget must make requests wait"
    mix = loop.checksums().copy()
    admit = loop.d_admit.cpu().numpy()[: len(trace.arrival)]
    assert (admit >= trace.arrival).all()
    for i in (0, len(trace.arrival) // 2, len(trace.arrival) - 1):
        alone = serve.Trace(trace.arrival[i:i + 1] * 0, trace.length[i:i + 1], trace.ids[i:i + 1])
        _run_all(loop, alone)
        assert int(loop.checksums()[0]) == int(mix[i]) == serve.expected_checksum(int(trace.ids[i]),
                                                                                   int(trace.length[i]))
    loop.close()
    a.close()


def test_serve_rejects_overflowing_config_and_bad_traces(mmc, cuda):
    """mmc_serve_create refuses a configuration whose 64-request admission sum
    could pass 2^32; mmc_serve_load refuses (on the device, without the host
    check) a trace with a zero or over-long length or a decreasing arrival,
    names the request and admits nothing.  A good trace then runs."""
    import ctypes
    from mallocmc_amd import serve
    a = _alloc(mmc)
    L = serve._bind()
    cfg = serve.MmcServeConfig(64, 1 << 20, 16, 256, 8192, 0, 64 * MiB)  # 64 x 512 x 1 MiB > 2^32
    h = ctypes.c_void_p()
    assert L.mmc_serve_create(a.handle, ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"2^32" in mmc.lib().mmc_last_error()
    loop = serve.ServeLoop(a, max_slots=64, max_tokens=100, budget_bytes=64 * MiB)
    ids = np.arange(4, dtype=np.uint32)
    for arrival, length, first in (([0, 0, 1, 1], [5, 0, 5, 5], 1), ([0, 0, 1, 1], [5, 5, 5, 101], 3),
                                   ([0, 2, 1, 3], [5, 5, 5, 5], 2)):
        t = serve.Trace(np.array(arrival, np.uint32), np.array(length, np.uint32), ids)
        with pytest.raises(mmc.MallocMCError, match=f"request {first}:"):
            loop.load(t, check=False)
        assert loop.stats()["admitted"] == 0
    good = serve.Trace(np.array([0, 0, 1, 1], np.uint32), np.array([5, 20, 1, 100], np.uint32), ids)
    st = _run_all(loop, good, chunk=64)
    assert st["completed"] == 4 and st["nulls"] == 0
    cs = loop.checksums()
    assert all(int(cs[i]) == serve.expected_checksum(int(ids[i]), int(good.length[i])) for i in range(4))
    loop.close()
    a.close()
