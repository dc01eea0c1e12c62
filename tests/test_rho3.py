"""Three-site RDM record (the reference's mipt.exe output).

CPU: the argument contract of cpp/apps/mipt.cpp:309-366.  GPU: the bulk
three-site RDM kernel against the FP64 oracle's RDMs (within 1e-5, with
trace 1 and Hermiticity), and the MIPTRHO2 file the driver writes
(cpp/include/mipt/density.hpp:10-48) parsed field by field, every record's
matrices against the oracle run of the same trajectory.
"""
import struct

import numpy as np
import pytest

from oracle import oracle as O


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


@pytest.mark.parametrize("argv,code", [
    (["--help"], 0),
    (["1"] * 9, 2),
    (["3"], 1),                                    # d must be 1 or 2
    (["1", "7"], 1),                               # odd ring for Haar brickwork
    (["1", "2"], 1),                               # fewer than three qubits
    (["1", "8", "-1"], 1),                         # negative periods
    (["1", "8", "2", "0"], 1),                     # realizations
    (["1", "8", "2", "2", "2", "0.5", "0.2"], 1),  # p_min > p_max
    (["1", "8", "2", "2", "2", "0", "1", "0"], 1),  # MMS is not built here
    (["2", "6"], 1),                               # 36 qubits
])
def test_mipt_argument_contract(M, argv, code, capfd):
    assert M.mipt_main(argv) == code
    if code == 1:
        assert "error:" in capfd.readouterr().err


def ring_triples(n):
    return [(i, (i + 1) % n, (i + 2) % n) for i in range(n)]


def grid_triples(n):
    out = [(i * n + j, i * n + (j + 1) % n, i * n + (j + 2) % n) for i in range(n) for j in range(n)]
    out += [(j * n + i, ((j + 1) % n) * n + i, ((j + 2) % n) * n + i) for i in range(n) for j in range(n)]
    return out


@pytest.mark.gpu
def test_batch_rho3_matches_oracle(M):
    cases = [(8, 8, 0.2, 301), (16, 6, 0.1, 302), (20, 4, 0.2, 303)]
    print("batch rho3", flush=True)
    for (n, layers, p, seed) in cases:
        b = M.Batch([M.Spec(n, layers, p, seed)])
        b.generate()
        b.evolve()
        b.synchronize()
        subs = ring_triples(n)
        rho = b.rho3(subs)[0]
        psi, *_ = O.simulate(n, layers, p, seed)
        for s, q in enumerate(subs):
            want = O.rdm(psi, n, list(q))
            assert np.abs(rho[s] - want).max() < 1e-5, (n, s)
            assert abs(np.trace(rho[s]) - 1) < 1e-12
            assert np.abs(rho[s] - rho[s].conj().T).max() < 1e-15
        b.close()
    # a grid and its row/column triples
    b = M.Batch([M.Spec(16, 8, 0.15, 304, True, grid_x=4)])
    b.generate()
    b.evolve()
    b.synchronize()
    subs = grid_triples(4)
    rho = b.rho3(subs)[0]
    psi, *_ = O.simulate_grid(4, 4, 8, 0.15, 304, True)
    for s, q in enumerate(subs):
        assert np.abs(rho[s] - O.rdm(psi, 16, list(q))).max() < 1e-5


def read_rho_file(path):
    data = open(path, "rb").read()
    assert data[:8] == b"MIPTRHO2"
    o = 8
    endian, version, dim, total, kept, mdim, periods, real, res, gx, gy, nsub = struct.unpack_from("<12I", data, o)
    o += 48
    (records,) = struct.unpack_from("<Q", data, o)
    o += 8
    p_min, p_max = struct.unpack_from("<2d", data, o)
    o += 16
    subs = np.frombuffer(data, np.uint32, nsub * kept, o).reshape(nsub, kept)
    o += 4 * nsub * kept
    head = dict(endian=endian, version=version, dim=dim, total=total, kept=kept, mdim=mdim, periods=periods,
                real=real, res=res, gx=gx, gy=gy, nsub=nsub, records=records, p_min=p_min, p_max=p_max)
    recs = []
    for _ in range(records):
        pi, r = struct.unpack_from("<2I", data, o)
        (p,) = struct.unpack_from("<d", data, o + 8)
        o += 16
        v = np.frombuffer(data, np.float64, nsub * mdim * mdim * 2, o).reshape(nsub, mdim, mdim, 2)
        o += 8 * v.size
        recs.append((pi, r, p, v[..., 0] + 1j * v[..., 1]))
    assert o == len(data)
    return head, subs, recs


def write_rho_file(path, head, subs, recs):
    """The layout read_rho_file parses, written back (used to pin that layout
    against the reference's own reader on the CPU)."""
    out = bytearray(b"MIPTRHO2")
    out += struct.pack("<12I", head["endian"], head["version"], head["dim"], head["total"], head["kept"], head["mdim"],
                       head["periods"], head["real"], head["res"], head["gx"], head["gy"], head["nsub"])
    out += struct.pack("<Q", head["records"]) + struct.pack("<2d", head["p_min"], head["p_max"])
    out += np.asarray(subs, np.uint32).tobytes()
    for (pi, r, p, rho) in recs:
        out += struct.pack("<2I", pi, r) + struct.pack("<d", p)
        out += np.stack([rho.real, rho.imag], -1).astype(np.float64).tobytes()
    open(path, "wb").write(bytes(out))


@pytest.mark.skipif(not O.ref_available(), reason="reference tree absent (oracle/_ref not built)")
def test_rho_file_layout_matches_reference_reader(tmp_path):
    # the layout this file's parser assumes, read back by the reference's
    # DensityMatrixReader (cpp/include/mipt/density.hpp:253-347)
    rng = np.random.default_rng(5)
    n = 6
    subs = ring_triples(n)
    recs = []
    for idx in range(4):
        psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
        psi /= np.linalg.norm(psi)
        recs.append((idx // 2, idx % 2, [0.1, 0.3][idx // 2], np.stack([O.rdm(psi, n, list(q)) for q in subs])))
    head = dict(endian=0x01020304, version=2, dim=1, total=n, kept=3, mdim=8, periods=2, real=2, res=2, gx=n, gy=1,
                nsub=n, records=4, p_min=0.1, p_max=0.3)
    path = str(tmp_path / "rho3.bin")
    write_rho_file(path, head, subs, recs)
    got = O.ref_density_read(path)
    assert (got["spatial_dimension"], got["total_qubits"], got["kept_qubits"], got["matrix_dimension"]) == (1, n, 3, 8)
    assert (got["realizations"], got["resolution"], got["record_count"], got["subsystem_count"]) == (2, 2, 4, n)
    assert [tuple(int(x) for x in s) for s in got["subsystems"]] == subs
    for i, (pi, r, p, rho) in enumerate(recs):
        assert (got["p_index"][i], got["realization"][i], got["p"][i]) == (pi, r, p)
        assert np.array_equal(got["rho"][i], rho)
    # a truncated file is rejected by the reference reader
    open(path, "r+b").truncate(len(open(path, "rb").read()) - 8)
    with pytest.raises(ValueError, match="incomplete"):
        O.ref_density_read(path)


@pytest.mark.gpu
@pytest.mark.parametrize("d,n", [(1, 6), (2, 3)])
def test_mipt_rho3_file_matches_oracle(M, tmp_path, monkeypatch, d, n):
    monkeypatch.chdir(tmp_path)
    monkeypatch.setenv("MIPT_SEED", "4242")
    periods, real, res = 2, 2, 2
    print(f"mipt d={d} n={n}", flush=True)
    assert M.mipt_main([d, n, periods, real, res, 0.1, 0.3, 1]) == 0
    fname = tmp_path / ("rho3.bin" if d == 1 else "rho3_2d.bin")
    head, subs, recs = read_rho_file(fname)
    if O.ref_available():
        # the driver's file read by the reference's own DensityMatrixReader
        ref = O.ref_density_read(str(fname))
        assert ref["record_count"] == head["records"] and ref["subsystem_count"] == head["nsub"]
        assert (ref["grid_x"], ref["grid_y"], ref["periods"]) == (head["gx"], head["gy"], head["periods"])
        for i, (pi, r, p, rho) in enumerate(recs):
            assert (ref["p_index"][i], ref["realization"][i], ref["p"][i]) == (pi, r, p)
            assert np.array_equal(ref["rho"][i], rho)
    L = n if d == 1 else n * n
    assert head == dict(endian=0x01020304, version=2, dim=d, total=L, kept=3, mdim=8, periods=periods, real=real,
                        res=res, gx=n, gy=1 if d == 1 else n, nsub=len(subs), records=real * res, p_min=0.1, p_max=0.3)
    want_subs = ring_triples(n) if d == 1 else grid_triples(n)
    assert [tuple(int(x) for x in s) for s in subs] == want_subs
    info = (tmp_path / "info.csv").read_text().splitlines()
    assert info[0].startswith("n,periods,realizations,resolution,p_min,p_max,circ_type")
    ps = [0.1, 0.3]
    for idx, (pi, r, p, rho) in enumerate(recs):
        assert (pi, r, p) == (idx // real, idx % real, ps[idx // real])
        seed = O.lib().orc_trajectory_seed(4242, idx)
        if d == 1:
            psi, *_ = O.simulate(L, 2 * periods, p, seed)
        else:
            psi, *_ = O.simulate_grid(n, n, 2 * periods, p, seed, True)
        for s, q in enumerate(want_subs):
            assert np.abs(rho[s] - O.rdm(psi, L, list(q))).max() < 1e-5
