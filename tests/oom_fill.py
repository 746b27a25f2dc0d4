"""Fill-until-first-null on the oracle (the reference's serialized protocol):
the fill fraction the reference's algorithm reaches on the config-4 size mix,
used as the yardstick for the GPU stress cycles (tests only)."""
import numpy as np


def oracle_first_null_fill(oracle_mod, heap_bytes, cycles=3, pagesize=65536, seed=1, regionsize=16):
    """Per cycle: allocate 80% U[16,256] B / 20% U[8 KiB, 60 KiB] requests
    until the first null, record payload bytes / page bytes and the free
    pages left, free a seeded random half.  Returns a list of dicts."""
    cfg = oracle_mod.make_config(pagesize=pagesize, regionsize=regionsize, accessblocksize=min(2 * 1024 ** 3, heap_bytes),
                                 resetfreedpages=1)
    out = []
    with oracle_mod.OracleHeap(cfg, heap_bytes) as h:
        g = h.geometry
        rng = np.random.default_rng(seed)
        live = []
        for _ in range(cycles):
            while True:
                b = int(rng.integers(16, 257)) if rng.random() < 0.8 else int(rng.integers(8192, 61441))
                o = h.malloc(b)
                if o is None:
                    break
                live.append((o, b))
            used = sum(((b + 15) // 16) * 16 for _, b in live)
            out.append({"fill_fraction": used / (g.numpages * pagesize), "free_pages": h.available_slots(pagesize),
                        "live": len(live)})
            keep = []
            for x in live:
                if rng.random() < 0.5:
                    h.free(x[0])
                else:
                    keep.append(x)
            live = keep
    return out
