"""Measures the tile choices of batch plans on the GPU and writes the tuning table the
library compiles in (tensor-blas_amd/csrc/batch_tuning.inc), keyed by (tier, op, size
bucket): the tile shape of SYRK and TRMM instances (64 x 256 or 128 x 128) and the waves
per workgroup of plans whose TRSM work lies in a bucket.  Each choice is what the
library's own autotuner (TB_BATCH_AUTOTUNE) picks on a plan of random instances of that
bucket, the faster of two timed runs per candidate.

    python tensor-blas_amd/python/tune_batch.py [--out PATH]   (then rebuild the library)
"""
import argparse
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

BUCKETS = [(64, 128), (128, 256), (256, 512), (512, 1024), (1024, 2048), (2048, 4096)]
FLOPS_PER_BUCKET = 4e10        # about this much work per (op, bucket) group


def sizes(g, lo, hi):
    n_mid = (lo * hi) ** 0.5
    count = int(min(96, max(4, FLOPS_PER_BUCKET / n_mid ** 3)))
    return np.rint(np.exp(g.uniform(np.log(lo), np.log(hi - 1), count))).astype(int).tolist()


def main():
    import torch
    import tensorblas

    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "..", "csrc", "batch_tuning.inc"))
    args = ap.parse_args()
    g = np.random.default_rng(0xC0FA)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(0xC0FA)

    def rnd(*shape):
        return torch.rand(*shape, device="cuda", generator=gen) * 2 - 1

    h = tensorblas.Handle()
    shapes, waves = {}, {}
    for tier in ("plain", "split"):
        # SYRK and TRMM instances of every bucket in one autotuned plan: the autotuner
        # times each (op, bucket) group on its own
        keep, entries = [], []
        for lo, hi in BUCKETS:
            for n in sizes(g, lo, hi):
                A, C = rnd(n, n), torch.zeros(n, n, device="cuda")
                keep += [A, C]
                entries.append(dict(op="syrk", n=n, k=n, A=A.data_ptr(), lda=n, C=C.data_ptr(), ldc=n))
            for n in sizes(g, lo, hi):
                L, B, C = rnd(n, n), rnd(n, n), torch.zeros(n, n, device="cuda")
                keep += [L, B, C]
                entries.append(dict(op="trmm", m=n, n=n, A=L.data_ptr(), lda=n, B=B.data_ptr(), ldb=n,
                                    C=C.data_ptr(), ldc=n, alpha=1.0))
        plan = tensorblas.Batch(h, tier, entries, autotune=True)
        t = plan.tiles()
        shapes[tier] = (t["syrk"], t["trmm"])
        print(f"{tier}: syrk {t['syrk']} trmm {t['trmm']} ({plan.info()['tune_us'] / 1e3:.0f} ms)",
              flush=True)
        plan.close()
        del keep
        torch.cuda.empty_cache()
        # waves: one autotuned plan of TRSM instances per bucket (2-plane plans only)
        waves[tier] = []
        for lo, hi in BUCKETS:
            if tier == "plain":
                waves[tier].append(4)
                continue
            keep, entries = [], []
            for n in sizes(g, lo, hi):
                L, B = rnd(n, n), rnd(n, n)
                L.view(n, n).diagonal().add_(float(n))
                keep += [L, B]
                entries.append(dict(op="trsm", m=n, n=n, A=L.data_ptr(), lda=n, C=B.data_ptr(), ldc=n))
            plan = tensorblas.Batch(h, tier, entries, autotune=True)
            waves[tier].append(plan.tiles()["waves"])
            plan.close()
            del keep
            torch.cuda.empty_cache()
        print(f"{tier}: waves {waves[tier]}", flush=True)
    h.close()
    dev = torch.cuda.get_device_properties(0)
    rows = []
    for tier in ("plain", "split"):
        sy, tr = shapes[tier]
        rows.append(f"    /* {tier} */ {{/* SYRK */ {{{', '.join(map(str, sy))}}}, "
                    f"/* TRMM */ {{{', '.join(map(str, tr))}}}}},")
    wrows = [f"    /* {tier} */ {{{', '.join(map(str, waves[tier]))}}}," for tier in ("plain", "split")]
    text = f"""// Tile choices of batch plans on MI355X, keyed by (tier, op, size bucket); included by
// host/tb_batch.cpp.  Buckets: n < 128, < 256, < 512, < 1024, < 2048, >= 2048 (SYRK n,
// TRMM m, TRSM n).  Shapes: 0 = 64 x 256 tiles, 1 = 128 x 128 tiles.  Waves: per
// workgroup of the persistent kernel (8 only for 2-plane plans).
// Written by tensor-blas_amd/python/tune_batch.py from tb_batch_tiles of autotuned plans.
// source: measured on {dev.name} ({dev.gcnArchName}, {dev.multi_processor_count} CUs)
static const int kTunedShape[2][kTunedOps][kBuckets] = {{
{chr(10).join(rows)}
}};
static const int kTunedWaves[2][kBuckets] = {{
{chr(10).join(wrows)}
}};
"""
    with open(args.out, "w") as f:
        f.write(text)
    print(f"wrote {os.path.normpath(args.out)}", flush=True)


if __name__ == "__main__":
    main()
