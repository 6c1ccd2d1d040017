"""Device plumbing for the GPU tests, over the library's own HIP runtime."""
import numpy as np


def gpu_available() -> bool:
    import ntt_amd
    return ntt_amd.device.is_available()


def to_dev(a: np.ndarray):
    import ntt_amd
    return ntt_amd.DeviceBuffer.from_numpy(np.ascontiguousarray(a, dtype=np.uint32))


def to_host(buf) -> np.ndarray:
    return buf.numpy(np.uint32)


def sync():
    import ntt_amd
    ntt_amd.synchronize()


def random_operand(logN: int, moduli, seed: int) -> np.ndarray:
    """Uniform words in [0, q_l) per limb, (limbs, N) uint32, from numpy's PCG64."""
    rng = np.random.default_rng(seed)
    qs = np.array([int(m[0]) for m in moduli], dtype=np.uint64)[:, None]
    return (rng.integers(0, 1 << 63, (len(moduli), 1 << logN), dtype=np.uint64) % qs).astype(np.uint32)


class DropIn:
    """Device copies of the arrays the reference's main() uploads (v04:947-967)."""

    def __init__(self, logN, moduli, domain):
        import ntt_amd
        self.logN, self.moduli, self.domain = logN, moduli, domain
        t = ntt_amd.make_tables(logN, moduli, domain)
        self.tw = to_dev(t.twiddles)
        self.itw = to_dev(t.inv_twiddles)
        self.q = to_dev(t.moduli)
        self.q_inv = to_dev(t.q_inv)
        self.r2 = to_dev(t.r2)
        self.deg = to_dev(t.degree_inv)
        self.L = len(moduli)

    def forward(self, A, cfg=None):
        import ntt_amd
        if self.domain == 1:
            ntt_amd.launch_ntt(A, self.tw, self.q, self.q_inv, self.r2, self.L, self.logN, cfg)
        else:
            ntt_amd.launch_ntt_plain(A, self.tw, self.q, self.L, self.logN, cfg)

    def inverse(self, A, cfg=None):
        import ntt_amd
        if self.domain == 1:
            ntt_amd.launch_intt(A, self.itw, self.q, self.q_inv, self.deg, self.L, self.logN, cfg)
        else:
            ntt_amd.launch_intt_plain(A, self.itw, self.q, self.deg, self.L, self.logN, cfg)
