"""Shared set-up for the parity tests: a seeded specimen, its state, the device pipeline
(GPUFunctionality over libicflow_hip.so) and the same inputs for the CPU oracle."""
import numpy as np

from icflow_amd import _lib, specimens
from icflow_amd.gpu_functionality import GPUFunctionality, broad_phase

H = 1e-4            # time step of the three-point-bend beam (0.40 x 0.05 x 0.05 m, 15,360 tets)
PENALTY_K = 1e11 * (0.1 / 40) ** 2   # 1e11 N/m^3 times the platen cell area (N/m per pair)
DIST_EPS = 1e-15    # ModelPrms.DistanceEpsilon default


def small_specimen(cells=(6, 5, 4), size=(0.06, 0.05, 0.04), grains=6, seed=0, platens=True,
                   anchor_bottom=True):
    col = specimens.make_specimen(cells, size, grains, seed, platens=cells[:2] if platens else None)
    if anchor_bottom:
        col.anchored = np.zeros(sum(len(m.x0) for m in col.meshes), bool)
        x = col.meshes[0].x0
        col.anchored[:len(x)] = np.isclose(x[:, 2], 0.0)     # bottom face of the specimen
    return col.finalize()


def make_state(col, consts, seed=1, platen_dz_frac=0.0, cells_z=4, size_z=0.04, **kw):
    hz = size_z / cells_z
    return specimens.synthetic_state(col, consts, seed, platen_dz=platen_dz_frac * hz, **kw)


class Pipeline:
    """One Newton iteration's device-side calls, in ImplicitModel3.Step order
    (ImplicitModel3.cs:383-398), with the oracle inputs captured along the way."""

    def __init__(self, col, consts, state, precision=_lib.ICF_FP64, contact=False):
        self.col, self.consts, self.state = col, consts, state
        gf = self.gf = GPUFunctionality(0)
        gf.SetConstants(consts)
        gf.set_precision(precision)
        gf.TransferStaticDataToDevice(col)
        gf.TransferNodesToGPU(state)
        self.impacts = None
        if contact:
            self.pairs = broad_phase(gf, gf.c_dn, gf.nd_stride, gf.c_ie, gf.el_all_stride, len(col.surf))
            self.impacts = gf.NarrowPhaseCollisionDetection(self.pairs)
        gf.CreateStructure(self.impacts)
        gf.TransferPCSR(state)

    def assemble(self, h=H):
        gf = self.gf
        counts = gf.AssembleElemsAndCZs(h)
        if self.impacts is not None and len(self.impacts):
            gf.collisionResponse(h, PENALTY_K, DIST_EPS)
        vals, rhs = gf.TransferLinearSystemToHost()
        st = gf.TransferUpdatedStateToHost()
        return dict(vals=vals, rhs=rhs, F=st["F"], dcz=st.get("dcz"), icz=st.get("icz"), counts=counts)


def oracle_assemble(orc, p, h=H, contact=True):
    """The same iteration on the CPU oracle, from the pipeline's packed host buffers."""
    gf = p.gf
    orc.set_constants(p.consts)
    dn = gf.c_dn.copy()
    ie = gf.c_ie_pcsr.copy()
    A = np.zeros(gf.csr.nnz * 9)
    b = np.zeros(gf.csr.N * 3)
    orc.elem_force(ie, dn, h, A, b, len(p.col.elastic), gf.el_elastic_stride, gf.nd_stride)
    out = dict()
    if gf.cz_stride:
        icz, dcz = gf.c_icz.copy(), gf.c_dcz.copy()
        orc.cz_force(dcz, icz, dn, A, b, h, len(p.col.czs), gf.cz_stride, gf.nd_stride)
        out.update(dcz=dcz, icz=icz)
    if contact and p.impacts is not None and len(p.impacts):
        icr, cr = gf.pack_icr(p.impacts)
        orc.collision_response(icr, dn, h, A, b, len(p.impacts), cr, gf.nd_stride, PENALTY_K, DIST_EPS)
    ns = gf.nd_stride
    out.update(vals=A, rhs=b, F=dn[12 * ns:15 * ns])
    return out
