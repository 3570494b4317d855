"""Synthetic workloads of the reference's benchmark configurations, built on the GPU.

Signatures come from the engine's device generator (``sigops_gen_*_dev``: seeded, per-item
deterministic signing), so millions of inputs take milliseconds and stay resident in HBM.
Configurations (sizes from the benchmark definitions):
  * ``secp(curve, n)``: all-valid (or ``corrupt_ppm``-corrupted) recoveries;
  * ``ed25519(n, max_len)``: messages of uniform length 0..max_len, 10 % corrupted by default;
  * ``mixed(n)``: 50 % secp256k1, 30 % secp256r1, 20 % Ed25519, schemes randomly interleaved,
    Ed25519 message lengths log-uniform on 0..4096 (L = floor(exp(U ln 4097)) - 1), 5 % corrupted.
Tensors are torch uint8/int64 on the given device; the host-side expectations (which items are
corrupted, the generator's public keys) come back with them.
"""
import numpy as np
import torch


def _ptr(t):
    return t.data_ptr() if t is not None else None


def _stream():
    return torch.cuda.current_stream().cuda_stream


class Items:
    """Device arrays of one workload. For secp: sigs (n,64), aux = digests (n,32), pubs (n,64).
    For Ed25519: aux = public keys, msgs (bytes), msg_off (n+1). corrupted: (n,) uint8."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def secp(ctx, curve, n, seed=1, corrupt_ppm=0, device="cuda:0"):
    dev = torch.device(device)
    sigs = torch.empty((n, 64), dtype=torch.uint8, device=dev)
    digs = torch.empty((n, 32), dtype=torch.uint8, device=dev)
    pubs = torch.empty((n, 64), dtype=torch.uint8, device=dev)
    cor = torch.empty(n, dtype=torch.uint8, device=dev)
    ctx.gen_secp_dev(curve, seed, n, corrupt_ppm, _ptr(sigs), _ptr(digs), _ptr(pubs), _ptr(cor), _stream())
    return Items(kind="secp", curve=curve, n=n, sigs=sigs, aux=digs, pubs=pubs, corrupted=cor)


def msg_offsets(lens):
    off = np.zeros(len(lens) + 1, np.int64)
    off[1:] = np.cumsum(lens)
    return off


def uniform_lengths(n, max_len, seed):
    return np.random.default_rng(seed).integers(0, max_len + 1, n)


def loguniform_lengths(n, max_len, seed):
    u = np.random.default_rng(seed).random(n)
    return (np.floor(np.exp(u * np.log(max_len + 1))) - 1).astype(np.int64).clip(0, max_len)


def ed25519(ctx, n, lens, seed=1, corrupt_ppm=100000, device="cuda:0"):
    dev = torch.device(device)
    off = msg_offsets(lens)
    total = int(off[-1])
    # 4-byte aligned, 16 bytes of slack after the last message (hash reads use aligned words)
    msgs = torch.zeros(((total + 16 + 3) // 4) * 4, dtype=torch.uint8, device=dev)
    d_off = torch.from_numpy(off).to(dev)
    sigs = torch.empty((n, 64), dtype=torch.uint8, device=dev)
    pks = torch.empty((n, 32), dtype=torch.uint8, device=dev)
    cor = torch.empty(n, dtype=torch.uint8, device=dev)
    ctx.gen_ed25519_dev(seed, n, corrupt_ppm, _ptr(d_off), _ptr(msgs), _ptr(sigs), _ptr(pks), _ptr(cor), _stream())
    return Items(kind="ed25519", n=n, sigs=sigs, aux=pks, msgs=msgs, msg_off=d_off, msg_off_host=off, corrupted=cor)


MIX = (0.5, 0.3, 0.2)


def mixed(ctx, n, seed=1, corrupt_ppm=50000, max_len=4096, device="cuda:0"):
    """Mixed workload: returns Items with schemes (n,), sigs, aux, msgs, msg_off, and per-scheme
    index arrays ``idx[s]`` (positions of scheme s in the mix) plus the generator outputs."""
    dev = torch.device(device)
    counts = [int(round(n * MIX[0])), int(round(n * MIX[1]))]
    counts.append(n - counts[0] - counts[1])
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    idx = [np.sort(perm[: counts[0]]), np.sort(perm[counts[0]: counts[0] + counts[1]]),
           np.sort(perm[counts[0] + counts[1]:])]
    k1 = secp(ctx, 0, counts[0], seed=seed * 3 + 0, corrupt_ppm=corrupt_ppm, device=device)
    r1 = secp(ctx, 1, counts[1], seed=seed * 3 + 1, corrupt_ppm=corrupt_ppm, device=device)
    lens_ed = loguniform_lengths(counts[2], max_len, seed + 17)
    ed = ed25519(ctx, counts[2], lens_ed, seed=seed * 3 + 2, corrupt_ppm=corrupt_ppm, device=device)
    schemes = np.zeros(n, np.uint8)
    schemes[idx[1]] = 1
    schemes[idx[2]] = 2
    lens = np.zeros(n, np.int64)
    lens[idx[2]] = lens_ed
    off = msg_offsets(lens)
    total = int(off[-1])
    msgs = torch.zeros(((total + 16 + 3) // 4) * 4, dtype=torch.uint8, device=dev)
    # place each Ed25519 message at its offset in the mixed layout (one gather of byte positions)
    src_pos = np.concatenate([np.arange(s, e) for s, e in zip(ed.msg_off_host[:-1], ed.msg_off_host[1:])]) \
        if total else np.zeros(0, np.int64)
    dst_pos = np.concatenate([np.arange(off[i], off[i + 1]) for i in idx[2]]) if total else np.zeros(0, np.int64)
    if total:
        msgs[torch.from_numpy(dst_pos).to(dev)] = ed.msgs[torch.from_numpy(src_pos).to(dev)]
    sigs = torch.empty((n, 64), dtype=torch.uint8, device=dev)
    aux = torch.empty((n, 32), dtype=torch.uint8, device=dev)
    cor = torch.empty(n, dtype=torch.uint8, device=dev)
    d_idx = [torch.from_numpy(i.astype(np.int64)).to(dev) for i in idx]
    for it, di in zip((k1, r1, ed), d_idx):
        sigs[di] = it.sigs
        aux[di] = it.aux
        cor[di] = it.corrupted
    return Items(kind="mixed", n=n, schemes=torch.from_numpy(schemes).to(dev), schemes_host=schemes, sigs=sigs,
                 aux=aux, msgs=msgs, msg_off=torch.from_numpy(off).to(dev), msg_off_host=off, corrupted=cor,
                 idx=idx, d_idx=d_idx, parts=(k1, r1, ed), lens=lens)
