"""The library's host-side set-up on a machine without a GPU.

hamstr_create and hamstr_batch_create build every host table -- the chain
slot table with its ghost rules, the per-cell accumulation lists, the orphan
diagonal tables, the ADI phases and launch groups -- before they look for a
device, so this code runs here (and under the sanitizer build of
tests/test_sanitizers.py).  Valid meshes therefore end in HAMSTR_ERR_NODEV
(or a context, on a machine with a GPU), and meshes that break the chain
contract are refused with HAMSTR_ERR_MESH before any device call.
"""
import numpy as np
import pytest

from hamstr_oop_amd import chains, hip, meshgen


def _valid(make):
    try:
        s = make()
    except hip.HamstrError as e:
        assert e.code == hip.ERR_NODEV, str(e)
        return
    s.close()


MESHES = {
    "circle16x8": lambda: meshgen.circle_ogrid(16, 8),
    "naca40x16_tri6": lambda: meshgen.naca_ogrid(40, 16, tri_layers=6),
    "naca24x10_t10": lambda: meshgen.naca_ogrid(24, 10, t=0.10),
}


@pytest.mark.parametrize("adi", ["exact", "fast"])
@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("name", list(MESHES))
def test_tables_of_generated_meshes(name, order, lhs, adi):
    mesh = MESHES[name]()
    _valid(lambda: hip.Solver(mesh, 0.5, 2.0, order, 10.0, lhs=hip.LHS_FP32 if lhs == "fp32" else hip.LHS_FP64,
                              adi_mode=hip.ADI_EXACT if adi == "exact" else hip.ADI_FAST))


def test_tables_of_long_chains_and_batches():
    # 1100-row loops (longer than a workgroup) and a batch of uneven meshes
    _valid(lambda: hip.Solver(meshgen.circle_ogrid(1100, 6), 0.3, 0.0, 3, 10.0, adi_mode=hip.ADI_FAST))
    cases = [(meshgen.circle_ogrid(16, 8), 0.3, 0.0), (meshgen.naca_ogrid(40, 16, tri_layers=6), 0.6, 1.0),
             (meshgen.naca_ogrid(64, 16), 0.5, -2.0)]
    for lhs in (hip.LHS_FP64, hip.LHS_FP32):
        _valid(lambda: hip.Batch(cases, 3, 10.0, lhs=lhs, adi_mode=hip.ADI_FAST))
    _valid(lambda: hip.Batch(cases, 5, 10.0, msweep=2, adi_mode=hip.ADI_EXACT))


def test_tables_of_a_builder_chained_mesh():
    m = meshgen.naca_ogrid(48, 16)
    rng = np.random.default_rng(5)
    conn = []
    for c, row in enumerate(m.conn.tolist()):
        if c // m.nc >= 8 and rng.random() < 0.5:
            a, b, cc, d = row
            conn += [[a, b, cc, cc], [a, cc, d, d]]
        else:
            conn.append(row)
    built = chains.build_chains(m.x, np.array(conn, np.int32))
    chains.validate(built, 5)
    _valid(lambda: hip.Solver(built, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST))


class _Broken:
    """A mesh description with one rule of the chain contract broken."""

    def __init__(self, mesh, **kw):
        self.x = mesh.x
        self.faces = np.array(kw.get("faces", mesh.faces), np.int32)
        self.chain_start = np.array(kw.get("chain_start", mesh.chain_start), np.int32)
        self.chain_conn = np.array(kw.get("chain_conn", mesh.chain_conn), np.int32)
        self.orphans = np.array(kw.get("orphans", mesh.orphans), np.int32)
        self.n_cell = mesh.n_cell
        self.n_face = mesh.n_face
        self.n_chain = len(self.chain_start) - 1


def _refused(m):
    with pytest.raises(hip.HamstrError) as e:
        s = hip.Solver(m, 0.5, 2.0, 3, 10.0)
        s.close()
    assert e.value.code == hip.ERR_MESH, str(e.value)


def test_broken_meshes_are_refused_before_any_device_call():
    m = meshgen.naca_ogrid(24, 8, tri_layers=2)
    cc = m.chain_conn.copy()
    cc[5] = cc[1]                      # a face on two chains
    _refused(_Broken(m, chain_conn=cc))
    cc = m.chain_conn.copy()
    cc[0] = m.n_face + 3               # out of range
    _refused(_Broken(m, chain_conn=cc))
    st = m.chain_start.copy()
    st[0] = 1                          # must start at 0
    _refused(_Broken(m, chain_start=st))
    f = m.faces.copy()
    b = int(np.nonzero(f[:, 4] < 0)[0][0])
    f[b, 4] = -5                       # an inter-domain face
    _refused(_Broken(m, faces=f))
    orph = np.concatenate([m.orphans, [m.orphans[0]]])   # listed twice
    _refused(_Broken(m, orphans=orph))
    orph = np.concatenate([m.orphans, [m.chain_conn[3]]])  # an orphan on a chain
    _refused(_Broken(m, orphans=orph))
