"""Address arithmetic of the 16-bit pair kernel (mesobasic_amd/csrc/
lbm_fused_pair.h), restated on the CPU with its 32-bit unsigned wrap: every
gather, seam patch and store of every thread of every workgroup stays inside
the population allocation.  Regression test for a version that put the -1
element shift of ez = +1 into the 32-bit offset: at node 0 the offset wrapped
to 2^32 - 1 and the gather addressed 4 GiB past the plane."""
import numpy as np
import pytest

EX = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0]
EY = [0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 1, -1]
EZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, -1, 1, -1, 1]
ZP = [5, 9, 11, 16, 18]     # ez = +1: seam patch of z = 0 pairs
ZM = [6, 10, 12, 15, 17]    # ez = -1: seam patch of z = nz-2 pairs
M = 0xFFFFFFFF


def plane_stride(n, eb):
    # lbm_capi.cpp plane_stride()
    per = 256 // eb
    ps = (n + per - 1) // per * per
    if (ps * eb) % 4096 == 0:
        ps += 512 // eb
    return ps


def element_range(nx, ny, nz, TY, TZ, kind, base_shift_in_offset=False):
    """[min, max] element index touched (16-bit elements from the start of
    the allocation) over every pull, patch and store of a step."""
    n = nx * ny * nz
    ps = plane_stride(n, 2)
    PZ = TZ // 2
    NRING = 2 * (PZ + 2) + 2 * TY
    plane = (ny * nz) & M
    zsh = 0 if kind == 0 else 1
    keep = 0 if kind == 0 else M
    tid = np.arange(TY * PZ)
    lo, hi = 1 << 62, -1
    for y0 in range(0, ny, TY):
        for z0 in range(0, nz, TZ):
            ly, lz = tid // PZ, 2 * (tid % PZ)
            yi, zi = y0 + ly, z0 + lz
            r = tid[tid < NRING]
            row = np.where(r < PZ + 2, 0, np.where(r < 2 * (PZ + 2), TY + 1,
                           1 + np.where(r < 2 * (PZ + 2) + TY, r - 2 * (PZ + 2), r - 2 * (PZ + 2) - TY)))
            col = np.where(r < PZ + 2, 2 * r, np.where(r < 2 * (PZ + 2), 2 * (r - (PZ + 2)),
                           np.where(r < 2 * (PZ + 2) + TY, 0, TZ + 2)))
            ry, rz = (y0 + row - 1) % ny, (z0 + col - 2) % nz
            for xp in range(nx):
                for Y, Z in ((yi, zi), (ry, rz)):
                    c = ((xp * ny + Y) * nz + Z) & M
                    xm = keep & ((nx - 1) * plane if xp == 0 else (-plane) & M)
                    xq = keep & ((-(nx - 1) * plane) & M if xp == nx - 1 else plane)
                    ym = keep & np.where(Y == 0, (ny - 1) * nz, (-nz) & M)
                    yq = keep & np.where(Y == ny - 1, (-(ny - 1) * nz) & M, nz)
                    for q in range(19):
                        o = np.zeros_like(c)
                        o = o + (xq if -EX[q] == 1 else 0) + (xm if -EX[q] == -1 else 0)
                        o = o + (yq if -EY[q] == 1 else 0) + (ym if -EY[q] == -1 else 0)
                        up = (c + o) & M
                        if base_shift_in_offset:
                            up = (up - EZ[q] * zsh) & M          # the faulting form
                            base = q * ps
                        elif EZ[q] == -1:
                            up = (up + zsh) & M
                            base = q * ps
                        else:
                            base = q * ps - (zsh if EZ[q] == 1 else 0)
                        el = base + ((up * 2) & M) // 2
                        lo, hi = min(lo, int(el.min())), max(hi, int(el.max()) + 1)
                        if kind and q in ZP and (Z == 0).any():
                            u = ((c + o + nz) & M)[Z == 0]
                            e2 = q * ps - 1 + ((u * 2) & M) // 2
                            lo, hi = min(lo, int(e2.min())), max(hi, int(e2.max()))
                        if kind and q in ZM and (Z == nz - 2).any():
                            u = ((c + o - (nz - 2)) & M)[Z == nz - 2]
                            e2 = q * ps + ((u * 2) & M) // 2
                            lo, hi = min(lo, int(e2.min())), max(hi, int(e2.max()))
            cs = (np.arange(nx)[:, None] * plane + (yi * nz + zi)[None, :]) & M
            lo, hi = min(lo, int(cs.min())), max(hi, 18 * ps + int(cs.max()) + 1)
    return lo, hi, 19 * ps


@pytest.mark.parametrize("dims", [(5, 16, 128), (3, 8, 256), (4, 24, 128)])
@pytest.mark.parametrize("tile", [(8, 128), (4, 128)])
@pytest.mark.parametrize("kind", [0, 1])
def test_pair_kernel_accesses_stay_in_the_allocation(dims, tile, kind):
    if dims[1] % tile[0] or dims[2] % tile[1]:
        pytest.skip("tile does not fit")
    lo, hi, total = element_range(*dims, *tile, kind)
    assert 0 <= lo and hi < total, (lo, hi, total)


def test_offset_shift_form_is_caught():
    # the arithmetic that faulted: caught by the same check
    lo, hi, total = element_range(3, 8, 128, 8, 128, 1, base_shift_in_offset=True)
    assert hi >= total
