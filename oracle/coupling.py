"""CPU restatement of the DNS coupling around the radiation call (test
infrastructure only: imported by tests/ as the checker, never by the product
package).

fix_temp (code/fix_var.f:49-73): each rank zeroes a full-size FP64 buffer
Tbuf(0:imax+1, 0:jmax+1, 0:kmax+1), writes T = cnew*Tc/Tplus + Tc over its
slab [startc, endc], and MPI_Allreduce(SUM) over all ranks gives the field,
assigned to the real(4) TMC that mc_gpu_ reads.

fix_rad (code/fix_var.f:75-121): MPI_Allreduce(SUM) of the node's qr / vr
(real(4), only the GPU owner's are nonzero), then every rank keeps its k range:
divQ(i, j, k - startc(3) + 1) = Qbuf(i, j, k) / stefan / Tc**4 for i = 1..imax,
j = 1..jmax/p_row, k = startc(3)..endc(3), zero elsewhere; vcore likewise.

Arrays are numpy arrays in Fortran order with the Fortran index ranges shifted
to 0-based storage (index 0 is the Fortran lower bound).
"""
import numpy as np

STEFAN = 5.670373e-08   # code/param.txt:19 (real*8)


def rank_buffer(cnew, startc, endc, n, Tc, Tplus):
    """One rank's Tbuf (code/fix_var.f:61-69): zeros outside its slab, FP64."""
    buf = np.zeros((n[0] + 2, n[1] + 2, n[2] + 2), np.float64, order="F")
    for k in range(startc[2], endc[2] + 1):
        for j in range(startc[1], endc[1] + 1):
            for i in range(startc[0], endc[0] + 1):
                c = cnew[i - startc[0] + 1, j - startc[1], k - startc[2]]   # cnew(i-startc(1)+1, j-startc(2)+1, ...)
                buf[i, j, k] = c * Tc / Tplus + Tc
    return buf


def allreduce_sum(buffers):
    """MPI_SUM over the ranks' buffers, in rank order (code/fix_var.f:71)."""
    out = np.zeros_like(buffers[0])
    for b in buffers:
        out = out + b
    return out


def fix_temp(slabs, n, Tc, Tplus):
    """TMC (real(4)) from the ranks' (cnew, startc, endc) (code/fix_var.f:49-73)."""
    return allreduce_sum([rank_buffer(c, s, e, n, Tc, Tplus) for c, s, e in slabs]).astype(np.float32)


def fix_rad(qr, startc, endc, Tc, n0, nj):
    """divQ of one rank from the node's (summed) qr(0:imax+1, 0:jmax/p_row+1,
    0:kmax+1) (code/fix_var.f:94-104); the same arithmetic gives vcore from vr."""
    nk = endc[2] - startc[2] + 1
    out = np.zeros((n0 + 2, nj, nk), np.float64, order="F")
    tc4 = Tc * Tc * Tc * Tc
    for k in range(startc[2], endc[2] + 1):
        for j in range(1, nj + 1):
            for i in range(1, n0 + 1):
                out[i, j - 1, k - startc[2]] = float(qr[i, j, k]) / STEFAN / tc4
    return out
