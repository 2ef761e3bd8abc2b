"""AMReX plotfile and VisMF files of MultiFabs in HBM (ksgf_plotfile_*,
ksgf_vismf_*, ksgf_checkpoint_write_header).

The text layouts are checked with readers written here after the
reference's own parsers: the plotfile Header in the field order of
tools/amrex_plotfile.py (parse_header), the checkpoint Header in the order of
tools/utilities/CheckpointToCSV/CheckpointFileUtil.cpp:40-120, and the FAB
payloads straight from the data file with numpy (the FAB header line's box,
component count and IEEE descriptor, then Fortran-order data, component
slowest).  The reader is checked on files written here independently (other
boxes, big-endian float32), and every round trip is bit for bit."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
CELL, XFACE = 0, 1


def _layout(lay, n_cell=(40, 24, 20), seed=3):
    geom = lay.Geometry(n_cell, prob_lo=(-1.0, 0.5, 2.0), prob_hi=(3.0, 2.9, 4.0), periodic=(1, 1, 0))
    rng = np.random.default_rng(seed)
    return lay.BoxArray(geom, lay.chop_random(n_cell, (8, 12, 16), rng, granularity=4))


def _box(text):
    t = text.replace("(", " ").replace(")", " ").replace(",", " ").split()
    v = [int(x) for x in t[:9]]
    return v[0:3], v[3:6], v[6:9]


def _parse_plot_header(path):
    """plotfile Header, fields in the order tools/amrex_plotfile.py reads them"""
    h = {}
    with open(path) as f:
        h["version"] = f.readline().rstrip("\n")
        h["ncomp"] = int(f.readline())
        h["names"] = [f.readline().rstrip() for _ in range(h["ncomp"])]
        h["spacedim"] = int(f.readline())
        h["time"] = float(f.readline())
        h["finest_level"] = int(f.readline())
        h["prob_lo"] = [float(x) for x in f.readline().split()]
        h["prob_hi"] = [float(x) for x in f.readline().split()]
        h["ref_ratio"] = [int(x) for x in f.readline().split()]
        h["domain"] = _box(f.readline())
        h["level_steps"] = [int(x) for x in f.readline().split()]
        h["dx"] = [float(x) for x in f.readline().split()]
        h["coord"] = int(f.readline())
        h["bwidth"] = int(f.readline())
        lev, ngrids, t = f.readline().split()
        h["lev"], h["ngrids"], h["lev_time"] = int(lev), int(ngrids), float(t)
        h["lev_step"] = int(f.readline())
        h["grids"] = [[[float(x) for x in f.readline().split()] for _ in range(3)] for _ in range(h["ngrids"])]
        h["mf"] = f.readline().rstrip()
    return h


def _read_vismf(prefix):
    """(boxes, per-FAB (box, ncomp, data (ncomp, nz, ny, nx))) read with numpy"""
    with open(prefix + "_H") as f:
        lines = f.read().split("\n")
    assert lines[0] == "1" and lines[1] == "1"
    ncomp, ngrow = int(lines[2]), int(lines[3])
    nb = int(lines[4].lstrip("(").split()[0])
    boxes = [_box(lines[5 + b]) for b in range(nb)]
    assert lines[5 + nb] == ")" and int(lines[6 + nb]) == nb
    fabs = []
    d = os.path.dirname(prefix)
    for b in range(nb):
        tag, name, off = lines[7 + nb + b].split()
        assert tag == "FabOnDisk:"
        with open(os.path.join(d, name), "rb") as fh:
            fh.seek(int(off))
            hl = fh.readline().decode()
            assert hl.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"), hl
            lo, hi, t = _box(hl[hl.index(")))") + 3:])
            nc = int(hl.rsplit(None, 1)[1])
            shape = (nc, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
            data = np.frombuffer(fh.read(8 * int(np.prod(shape))), dtype="<f8").reshape(shape)
        fabs.append(((lo, hi, t), nc, data))
    return ncomp, ngrow, boxes, fabs


def _rand_mf(dba, ba, itype, ncomp, ngrow, seed):
    rng = np.random.default_rng(seed)
    h = rng.standard_normal(ba.mf_size(itype, ncomp, ngrow))
    return h, dba.upload(h)


def test_plotfile_round_trip_and_header(ksgf, gpu, tmp_path):
    lay = ksgf.layout
    ba = _layout(lay)
    dba = ksgf.DeviceBoxArray(ba)
    names = ["velocityx", "velocityy", "velocityz"]
    h, mf = _rand_mf(dba, ba, CELL, 3, 2, 11)
    print(f"plotfile of {ba.nbox} boxes", flush=True)
    dba.write_plotfile(tmp_path / "plt00007", mf, 3, 2, names, time=0.125, step=7)
    H = _parse_plot_header(tmp_path / "plt00007" / "Header")
    g = ba.geom
    assert H["version"] == "HyperCLaw-V1.1" and H["ncomp"] == 3 and H["names"] == names
    assert H["spacedim"] == 3 and H["time"] == 0.125 and H["finest_level"] == 0
    assert H["prob_lo"] == list(g.prob_lo) and H["prob_hi"] == list(g.prob_hi) and H["ref_ratio"] == []
    assert H["domain"] == ([0, 0, 0], list(g.dhi), [0, 0, 0])
    assert H["level_steps"] == [7] and H["dx"] == list(g.dx) and H["coord"] == 0 and H["bwidth"] == 0
    assert (H["lev"], H["ngrids"], H["lev_time"], H["lev_step"]) == (0, ba.nbox, 0.125, 7)
    for b in range(ba.nbox):
        for d in range(3):
            lo = g.prob_lo[d] + g.dx[d] * ba.boxes[b, d]
            hi = g.prob_lo[d] + g.dx[d] * (ba.boxes[b, 3 + d] + 1)
            assert H["grids"][b][d] == pytest.approx([lo, hi], rel=1e-15, abs=1e-15)
    assert H["mf"] == "Level_0/Cell"
    # payload: the valid cells of every FAB, in box order
    ncomp, ngrow, boxes, fabs = _read_vismf(str(tmp_path / "plt00007" / "Level_0" / "Cell"))
    assert (ncomp, ngrow, len(boxes)) == (3, 0, ba.nbox)
    for b in range(ba.nbox):
        v = ba.fab_view(h, b, CELL, 3, 2)[:, 2:-2, 2:-2, 2:-2]
        (lo, hi, t), nc, data = fabs[b]
        assert lo == list(ba.boxes[b, :3]) and hi == list(ba.boxes[b, 3:]) and t == [0, 0, 0] and nc == 3
        assert np.array_equal(data, v)
    # the header's per-FAB minima and maxima
    with open(tmp_path / "plt00007" / "Level_0" / "Cell_H") as f:
        txt = f.read().split("\n")
    i = txt.index(f"{ba.nbox},3")
    mins = [[float(x) for x in txt[i + 1 + b].rstrip(",").split(",")] for b in range(ba.nbox)]
    for b in range(ba.nbox):
        v = ba.fab_view(h, b, CELL, 3, 2)[:, 2:-2, 2:-2, 2:-2]
        assert mins[b] == [float(v[n].min()) for n in range(3)]
    # read back into a MultiFab with one ghost layer: valid cells equal, ghosts untouched
    out = dba.alloc(CELL, 3, 1)
    dba.read_plotfile(tmp_path / "plt00007", out, 3, 1)
    o = out.cpu().numpy()
    for b in range(ba.nbox):
        assert np.array_equal(ba.fab_view(o, b, CELL, 3, 1)[:, 1:-1, 1:-1, 1:-1],
                              ba.fab_view(h, b, CELL, 3, 2)[:, 2:-2, 2:-2, 2:-2])
    assert np.count_nonzero(o) == 3 * int(np.prod(g.n_cell))
    dba.close()


def test_vismf_checkpoint_round_trip_with_ghosts(ksgf, gpu, tmp_path):
    lay = ksgf.layout
    ba = _layout(lay, seed=5)
    dba = ksgf.DeviceBoxArray(ba)
    dba.write_checkpoint_header(tmp_path / "chk00012", 12, 3.5, 0.01, 0.011, 0.012)
    with open(tmp_path / "chk00012" / "Header") as f:
        lines = f.read().split("\n")
    # field order of CheckpointFileUtil.cpp:40-120
    assert lines[0] == "Checkpoint version: 1" and int(lines[1]) == 0 and int(lines[2]) == 12
    assert [float(x) for x in lines[3:7]] == [3.5, 0.01, 0.011, 0.012]
    assert [float(x) for x in lines[7].split()] == list(ba.geom.prob_lo)
    assert [float(x) for x in lines[8].split()] == list(ba.geom.prob_hi)
    assert lines[9] == f"({ba.nbox} 0"
    for b in range(ba.nbox):
        lo, hi, t = _box(lines[10 + b])
        assert lo == list(ba.boxes[b, :3]) and hi == list(ba.boxes[b, 3:]) and t == [0, 0, 0]
    assert lines[10 + ba.nbox] == ")"
    for itype, ncomp, ngrow in ((CELL, 3, 3), (XFACE + 2, 1, 1)):
        h, mf = _rand_mf(dba, ba, itype, ncomp, ngrow, 20 + itype)
        prefix = tmp_path / "chk00012" / "Level_0" / f"f{itype}"
        print(f"VisMF itype {itype} ncomp {ncomp} ngrow {ngrow}", flush=True)
        dba.write_vismf(prefix, mf, itype, ncomp, ngrow)
        nc, ng, boxes, fabs = _read_vismf(str(prefix))
        assert (nc, ng) == (ncomp, ngrow)
        for b in range(ba.nbox):
            lo, hi, t = boxes[b]
            nodal = [int(itype == d + 1) for d in range(3)]
            assert t == nodal and lo == list(ba.boxes[b, :3]) and hi == list(ba.boxes[b, 3:] + np.array(nodal))
            assert np.array_equal(fabs[b][2], ba.fab_view(h, b, itype, ncomp, ngrow))
        out = dba.alloc(itype, ncomp, ngrow)
        dba.read_vismf(prefix, out, itype, ncomp, ngrow)
        assert np.array_equal(out.cpu().numpy(), h)
    dba.close()


def _write_foreign_plotfile(root, geom, boxes, field, names, big_endian_float):
    """a one-level plotfile written with numpy only (other boxes, float32)"""
    os.makedirs(os.path.join(root, "Level_0"), exist_ok=True)
    nc = field.shape[0]
    offs = []
    with open(os.path.join(root, "Level_0", "Cell_D_00000"), "wb") as f:
        for b in boxes:
            lo, hi = b[:3], b[3:]
            offs.append(f.tell())
            desc = "(4, (32 8 23 0 1 9 0 127)),(4, (1 2 3 4))" if big_endian_float else \
                   "(8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1))"
            f.write(f"FAB ({desc})(({lo[0]},{lo[1]},{lo[2]}) ({hi[0]},{hi[1]},{hi[2]}) (0,0,0)) {nc}\n".encode())
            blk = field[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
            f.write(np.ascontiguousarray(blk).astype(">f4" if big_endian_float else "<f8").tobytes())
    with open(os.path.join(root, "Level_0", "Cell_H"), "w") as f:
        f.write(f"1\n1\n{nc}\n0\n({len(boxes)} 0\n")
        for b in boxes:
            f.write(f"(({b[0]},{b[1]},{b[2]}) ({b[3]},{b[4]},{b[5]}) (0,0,0))\n")
        f.write(f")\n{len(boxes)}\n")
        for o in offs:
            f.write(f"FabOnDisk: Cell_D_00000 {o}\n")
        f.write("\n")
    dx = geom.dx
    with open(os.path.join(root, "Header"), "w") as f:
        f.write(f"HyperCLaw-V1.1\n{nc}\n" + "".join(n + "\n" for n in names) + "3\n0.5\n0\n")
        f.write(" ".join(repr(x) for x in geom.prob_lo) + " \n" + " ".join(repr(x) for x in geom.prob_hi) + " \n\n")
        hi = geom.dhi
        f.write(f"((0,0,0) ({hi[0]},{hi[1]},{hi[2]}) (0,0,0)) \n3 \n" + " ".join(repr(x) for x in dx) + " \n0\n0\n")
        f.write(f"0 {len(boxes)} 0.5\n3\n")
        for b in boxes:
            for d in range(3):
                f.write(f"{geom.prob_lo[d] + dx[d] * b[d]!r} {geom.prob_lo[d] + dx[d] * (b[3 + d] + 1)!r}\n")
        f.write("Level_0/Cell\n")


@pytest.mark.parametrize("big_endian_float", [False, True])
def test_read_plotfile_written_elsewhere(ksgf, gpu, tmp_path, big_endian_float):
    lay = ksgf.layout
    ba = _layout(lay, seed=9)
    dba = ksgf.DeviceBoxArray(ba)
    nx, ny, nz = ba.geom.n_cell
    rng = np.random.default_rng(4)
    field = rng.standard_normal((2, nz, ny, nx))
    boxes = [(0, 0, 0, nx // 2 - 1, ny - 1, nz - 1), (nx // 2, 0, 0, nx - 1, ny - 1, nz - 1)]
    root = str(tmp_path / "foreign")
    _write_foreign_plotfile(root, ba.geom, boxes, field, ["a", "b"], big_endian_float)
    out = dba.alloc(CELL, 2, 2)
    print(f"foreign plotfile, big-endian float32 = {big_endian_float}", flush=True)
    dba.read_plotfile(root, out, 2, 2)
    want = field.astype(np.float32).astype(np.float64) if big_endian_float else field
    assert np.array_equal(ba.gather(out.cpu().numpy(), CELL, 2, 2), want)
    with pytest.raises(ksgf.KsgfError):
        dba.read_plotfile(root, dba.alloc(CELL, 3, 0), 3, 0)
    dba.close()


def test_stepper_plotfile(ksgf, workloads, gpu, tmp_path):
    """the host stepper's plotfile (IOManager::write_plot_file) holds its
    velocity, gp and density after a step, and reads back into a MultiFab"""
    import importlib

    hs = importlib.import_module("kynema-sgf_amd.host_stepper")
    case = workloads.taylor_green(32, max_grid_size=16)
    S = hs.HostStepper(case)
    S.load()
    S.step()
    print("stepper plotfile", flush=True)
    S.write_plotfile(tmp_path / "plt00001", time=0.5, step=1)
    H = _parse_plot_header(tmp_path / "plt00001" / "Header")
    assert H["names"] == ["velocityx", "velocityy", "velocityz", "gpx", "gpy", "gpz", "density"]
    ba = case.ba
    dba = ksgf.DeviceBoxArray(ba)
    out = dba.alloc(CELL, 7, 0)
    dba.read_plotfile(tmp_path / "plt00001", out, 7, 0)
    f = ba.gather(out.cpu().numpy(), CELL, 7, 0)
    assert np.array_equal(f[:3], S.velocity())
    assert np.all(f[6] == 1.0)
    S.close()
    dba.close()
