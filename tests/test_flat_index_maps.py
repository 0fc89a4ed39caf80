"""CPU check of the indices the flat-box stage kernels form by arithmetic
instead of loading vmapM / vmapP: the element each thread of each workgroup
works on (xcd_remap, several elements per workgroup, padding elements in the
last workgroup) and the exterior node of every face node (neighbour element
and its opposite face node, or the own node at a closed boundary).

The same arithmetic as heve_kernel<N1, EPB, MODE_FUSED_FLAT> and heve7_body
(fe-project_amd/csrc/fedg_kernels.hip, fedg_device.h), restated here and
compared with the builder's maps, which the general kernels load:
  - every element below Ne is taken by exactly one (workgroup, slot);
  - padding slots (ke >= Ne) exist only in the workgroup that xcd_remap sends
    to the last block, and the kernels load nothing for them;
  - an open face's exterior node is vmapP - 1, resolved through the halo slot
    (halo_src) at a periodic boundary; a closed face's is the own node
    vmapM - 1, which the halo slot vmapP points to copies, with the boundary's
    velocity BC;
  - every index formed for a valid element is below Np * Ne (the fused
    stepper's fields have no halo tail).
Meshes cover p = 1..7, a number of elements that is and is not a multiple of
the elements per workgroup and of 8, and periodic and walled boundaries,
including the 5 x 3 x 4 p = 1 box of a removed tracer-kernel variant that once
faulted the GPU.
"""
import numpy as np
import pytest

import fedg

BND_SLIP = 1
OPP = (2, 3, 0, 1, 5, 4)


def xcd_remap(b, nb):
    """fedg_device.h xcd_remap: XCD x gets a contiguous 1/8 of the blocks."""
    if nb & 7:
        return b
    return (b & 7) * (nb >> 3) + (b >> 3)


def face_node(N1, f, a):
    """fedg_device.h face_node (faces y-, x+, y+, x-, z-, z+)."""
    N, n2 = N1 - 1, N1 * N1
    a0, a1 = a % N1, a // N1
    return (a0 + n2 * a1, N + N1 * a0 + n2 * a1, a0 + N1 * N + n2 * a1, N1 * a0 + n2 * a1, a, a + n2 * N)[f]


def exterior(mesh, ke, f, a):
    """Exterior node of face node a of face f of element ke, as the flat kernels form it."""
    N1, Np = mesh.porder + 1, mesh.Np
    NeX, NeY, NeZ = mesh.NeX, mesh.NeY, mesh.NeZ
    px, py, pz = mesh.periodic
    ei, ej, ek = ke % NeX, (ke // NeX) % NeY, ke // (NeX * NeY)
    ni, nj, nk = ei, ej, ek
    closed = False
    if f == 0:
        nj -= 1
        if nj < 0:
            nj, closed = (NeY - 1, False) if py else (nj, True)
    elif f == 1:
        ni += 1
        if ni >= NeX:
            ni, closed = (0, False) if px else (ni, True)
    elif f == 2:
        nj += 1
        if nj >= NeY:
            nj, closed = (0, False) if py else (nj, True)
    elif f == 3:
        ni -= 1
        if ni < 0:
            ni, closed = (NeX - 1, False) if px else (ni, True)
    elif f == 4:
        nk -= 1
        if nk < 0:
            nk, closed = (NeZ - 1, False) if pz else (nk, True)
    else:
        nk += 1
        if nk >= NeZ:
            nk, closed = (0, False) if pz else (nk, True)
    if closed:
        return Np * ke + face_node(N1, f, a), True
    return Np * (ni + NeX * (nj + NeY * nk)) + face_node(N1, OPP[f], a), False


def epb_for(N1):
    """Elements per workgroup of heve_kernel (fedg_kernels.hip epb_for)."""
    Np = N1 ** 3
    return 1 if Np >= 128 else 256 // Np


MESHES = [
    (1, 5, 3, 4, (True, True, False)),    # the removed variant's faulting case shape: 60 elements, EPB 32
    (1, 5, 3, 4, (False, True, False)),
    (1, 8, 8, 4, (True, True, False)),    # 256 elements, 8 blocks: xcd_remap permutes
    (2, 7, 2, 5, (False, False, False)),  # 70 elements, EPB 9: 8 blocks, 2 padding slots
    (3, 8, 8, 4, (True, True, False)),    # config 1 shape, EPB 4: 64 blocks
    (3, 5, 3, 3, (False, True, False)),   # 45 elements, EPB 4: 12 blocks, 3 padding
    (4, 4, 3, 4, (True, True, True)),
    (5, 3, 3, 3, (False, False, False)),
    (6, 5, 3, 4, (True, False, False)),
    (7, 4, 3, 4, (True, True, False)),    # 48 elements: heve7 grid a multiple of 8
    (7, 5, 3, 4, (False, True, False)),   # 60 elements: not a multiple of 8
]


@pytest.mark.parametrize("p,NeX,NeY,NeZ,periodic", MESHES)
def test_flat_kernel_indices_match_builder_maps(p, NeX, NeY, NeZ, periodic):
    mesh = fedg.BoxMesh(p, NeX, NeY, NeZ, (0.0, 1.0e3 * NeX), (0.0, 1.0e3 * NeY), (0.0, 1.0e3 * NeZ),
                        periodic=periodic)
    N1, Np, Nfp, NfpTot, Ne = p + 1, mesh.Np, mesh.Nfp_h, mesh.NfpTot, mesh.Ne
    assert Nfp == N1 * N1
    velbc = mesh.velbc([0 if periodic[a] else BND_SLIP for a in (1, 0, 1, 0, 2, 2)])
    vmapM = mesh.vmapM.reshape(Ne, NfpTot) - 1
    vmapP = mesh.vmapP.reshape(Ne, NfpTot) - 1
    # heve_kernel<N1, EPB, MODE_FUSED_FLAT>: nblk workgroups of EPB elements,
    # and the p = 7 heve7 kernels: Ne workgroups of one element
    layouts = [(epb_for(N1), -(-Ne // epb_for(N1)))]
    if N1 == 8:
        layouts.append((1, Ne))
    for EPB, nblk in layouts:
        taken = np.zeros(Ne, dtype=np.int64)
        padding_blocks = set()
        for b in range(nblk):
            blk = xcd_remap(b, nblk)
            assert 0 <= blk < nblk
            for e in range(EPB):
                ke = blk * EPB + e
                if ke >= Ne:
                    padding_blocks.add(blk)  # the kernels load and store nothing for it
                    continue
                taken[ke] += 1
        assert np.all(taken == 1), "every element exactly once"
        assert padding_blocks <= {nblk - 1}, "padding only in the last block"
        assert nblk * EPB - Ne < EPB
    # exterior node of every face node of every element
    for ke in range(Ne):
        for f in range(6):
            for a in range(Nfp):
                fp = f * Nfp + a
                iM = Np * ke + face_node(N1, f, a)
                assert iM == vmapM[ke, fp]
                iP, closed = exterior(mesh, ke, f, a)
                assert 0 <= iP < Np * Ne
                # vmapP points into the halo tail at the domain boundary (the
                # single-tile exchange fills the slot from halo_src)
                slot = vmapP[ke, fp] - Np * Ne
                src = mesh.halo_src[slot] - 1 if slot >= 0 else vmapP[ke, fp]
                if closed:
                    assert slot >= 0, "a closed face's vmapP points into the halo"
                    assert src == iM == iP
                    assert velbc[slot] == BND_SLIP
                else:
                    assert iP == src, (ke, f, a)
                    if slot >= 0:
                        assert velbc[slot] == 0, "a periodic face has no velocity BC"
