"""MFMA utilisation of a kernel from one rocprofv3 --pmc run with
SQ_VALU_MFMA_BUSY_CYCLES, GRBM_GUI_ACTIVE and SQ_INSTS_VALU_MFMA_MOPS_F32:
utilisation = busy cycles summed over SIMDs / (active cycles x SIMDs), the
rocprofv3 MfmaUtil expression; MFMA flops = MOPS_F32 x 512.

    python -m krans_amd.pmc_mfma <run_results.db> <kernel-substring> <simds> [--largest]

With --largest only the dispatches whose MFMA op count is at least half the
largest one count.
"""
import json
import sqlite3
import sys


def per_dispatch(db, counter, like):
    c = sqlite3.connect(db)
    return {r[0]: r[1] for r in c.execute(
        "select dispatch_id, sum(value) from counters_collection where counter_name = ? "
        "and kernel_name like ? group by dispatch_id", (counter, f"%{like}%"))}


def main(db, like, simds, largest):
    busy = per_dispatch(db, "SQ_VALU_MFMA_BUSY_CYCLES", like)
    act = per_dispatch(db, "GRBM_GUI_ACTIVE", like)
    mops = per_dispatch(db, "SQ_INSTS_VALU_MFMA_MOPS_F32", like)
    ids = sorted(set(busy) & set(act) & set(mops))
    if largest and ids:
        top = max(mops[i] for i in ids)
        ids = [i for i in ids if mops[i] >= 0.5 * top]
    util = [busy[i] / (act[i] * simds) for i in ids if act[i] > 0]
    out = {"kernel_like": like, "largest_only": largest, "dispatches": len(ids),
           "simds": simds,
           "mfma_util_mean": sum(util) / max(1, len(util)),
           "mfma_flops_per_dispatch": sum(mops[i] for i in ids) * 512.0 / max(1, len(ids)),
           "gpu_active_cycles_per_dispatch": sum(act[i] for i in ids) / max(1, len(ids))}
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2], int(sys.argv[3]), "--largest" in sys.argv[4:])
