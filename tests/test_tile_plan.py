"""Host-side checks of the batch's tile plans (no GPU): every sweep schedule
that `build_layer_plan` produces, read through `mipt_hip_layer_plan`.

A plan is correct when the data movement it asks of the sweep kernel is
well formed: each stage's register/lane/wave assignment is a permutation of
the tile's u bits, every bond of the layer is applied exactly once inside a
window holding both of its sites and on the register slots the kernel's bond
forms read (bond[0] on register bits (1,2) as (q0, q1), or (0,1) / (1,0) for
a bond of qubit 0; bond[1] on (3,4)), qubit 0 stays on register bit 0 (the
LDS exchanges split the tile on it), a stage without an exchange is the
load/store mapping, and the one stage holding the bond of qubit 0 comes last
(the sweep kernels fix its form at compile time and refuse other plans).
The bond lists are checked against the reference's brickwork
(`cpp/include/mipt/haar.hpp:233-243, 300-304`: pairs (2b + t mod 2, next)
with the closing (n-1, 0) bond on a ring).  Every sweep and fused pair of
the benchmarked chains has a kernel compiled for its stage structure.
"""
import itertools

import pytest

CONFIGS = {  # (tile_bits, threads) -> (register bits, wave bits)
    (13, 0): (5, 2),
    (14, 0): (6, 2),
    (13, 128): (6, 1),
    (12, 0): (6, 0),
}


@pytest.fixture(scope="module")
def M():
    import mipt_amd
    mipt_amd.lib()
    return mipt_amd


def m0_map(rb, wb):
    reg = [0] + [6 + r for r in range(1, rb)]
    lane = [b + 1 for b in range(6)]
    wave = [6 + rb + b for b in range(wb)]
    return reg, lane, wave


def check_plan(plan, L, nq, closed, gx, layer, rb, wb, tag):
    bonds = plan["bonds"]
    if gx == 0:
        # brickwork of the reference: start = layer mod 2, closing bond on a ring
        start = layer % 2
        want = [(q, q + 1) for q in range(start, L - 1, 2)]
        if closed and start == 1 and L % 2 == 0:
            want.append((L - 1, 0))
        assert bonds == want, f"{tag}: bonds {bonds}"
    seen = [0] * len(bonds)
    for pi, ps in enumerate(plan["passes"]):
        k, c, s = ps["k"], ps["c"], ps["s"]
        assert k == 6 + rb + wb, tag
        u = lambda q: q if q < c else (q - s + c if s <= q < s + (k - c) else None)  # noqa: E731
        nst = len(ps["stages"])
        zp_stages = [i for i, st in enumerate(ps["stages"]) if st["zpair"]]
        assert len(zp_stages) <= 1 and (not zp_stages or zp_stages[0] == nst - 1), f"{tag} pass {pi}: {zp_stages}"
        for si, st in enumerate(ps["stages"]):
            where = f"{tag} pass {pi} stage {si}"
            reg, lane, wave = st["reg_u"][:rb], st["lane_u"], st["wave_u"][:wb]
            assert sorted(reg + lane + wave) == list(range(k)), f"{where}: mapping {reg} {lane} {wave}"
            assert reg[0] == 0, f"{where}: u bit 0 on register bit {reg.index(0) if 0 in reg else None}"
            if not st["use_lds"]:
                assert (reg, lane, wave) == m0_map(rb, wb), f"{where}: exchange-free stage off the load mapping"
                assert st["zpair"] == 0, where
            for j, (b, q0, q1) in enumerate(st["bonds"]):
                if b < 0:
                    continue
                assert (q0, q1) == tuple(bonds[b]), f"{where}: bond {b}"
                seen[b] += 1
                u0, u1 = u(q0), u(q1)
                assert u0 is not None and u1 is not None, f"{where}: bond ({q0},{q1}) outside window s={s}"
                if j == 1:
                    assert (st["reg_u"][3], st["reg_u"][4]) == (u0, u1), f"{where}: bond[1] slots"
                elif st["zpair"] == 0:
                    assert (st["reg_u"][1], st["reg_u"][2]) == (u0, u1), f"{where}: bond[0] slots"
                elif st["zpair"] == 1:
                    assert u0 == 0 and st["reg_u"][1] == u1, f"{where}: zpair 1"
                else:
                    assert u1 == 0 and st["reg_u"][1] == u0, f"{where}: zpair 2"
            if st["zpair"]:
                b, q0, q1 = st["bonds"][0]
                assert b >= 0 and 0 in (q0, q1), where
    assert seen == [1] * len(bonds), f"{tag}: bonds applied {seen}"


GEOMETRIES = [(L, L, closed, 0) for L in range(15, 29) for closed in (0, 1)] + \
             [(L, L + 1, 1, 0) for L in (16, 20, 24)] + \
             [(L, L, closed, gx) for (L, gx) in ((16, 4), (20, 4), (20, 5), (24, 4), (24, 6), (25, 5))
              for closed in (0, 1)]


@pytest.mark.parametrize("cfg", sorted(CONFIGS))
def test_every_plan_is_well_formed(M, cfg):
    rb, wb = CONFIGS[cfg]
    k = 6 + rb + wb
    checked = 0
    for (L, nq, closed, gx), layer, plan in itertools.product(GEOMETRIES, range(4), range(4)):
        if nq <= k:
            continue
        tag = f"L={L} nq={nq} closed={closed} gx={gx} layer={layer} plan={plan} cfg={cfg}"
        try:
            p = M.layer_plan(L, nq, closed, gx, layer, cfg[0], cfg[1], plan)
        except M.MiptHipError:
            # a grid whose column bonds no window covers is refused, never mis-planned;
            # chains always have a plan
            assert gx > 0, tag
            continue
        check_plan(p, L, nq, closed, gx, layer, rb, wb, tag)
        checked += 1
    assert checked > 100


def test_l20_ring_uses_two_sweeps_of_at_most_three_stages(M):
    # the metric's workload: two windows (qubits 5.. and 12..), the bond of
    # qubit 0 in the last stage of its sweep
    for layer in (0, 1):
        p = M.layer_plan(20, 20, True, 0, layer)
        assert [ps["s"] for ps in p["passes"]] == [5, 12]
        assert all(len(ps["stages"]) <= 3 for ps in p["passes"])
        zp = [(pi, len(ps["stages"]) - 1) for pi, ps in enumerate(p["passes"]) if ps["stages"][-1]["zpair"]]
        assert len(zp) == 1


def test_layer_plan_refuses_bad_arguments(M):
    with pytest.raises(M.MiptHipError):
        M.layer_plan(20, 20, True, 0, 0, 13, 64)     # not a tile configuration
    with pytest.raises(M.MiptHipError):
        M.layer_plan(12, 12, True, 0, 0)             # fits one tile: the small kernel's case
    with pytest.raises(M.MiptHipError):
        M.layer_plan(20, 20, True, 3, 0)             # 3 does not divide 20



def structure_code(ps):
    """A sweep's stage-structure code as the compiled kernels key it: the
    stage count, plus per stage si the bits 1 << (2 + 3 si) (exchange),
    1 << (3 + 3 si) (bond[0]), 1 << (4 + 3 si) (bond[1]); None for a plan the
    codes do not describe.  Returns (qubit-0 form, code)."""
    st = ps["stages"]
    if not 1 <= len(st) <= 3:
        return None
    c = len(st)
    for si, x in enumerate(st):
        if (si > 0 and not x["use_lds"]) or (x["zpair"] and (si != len(st) - 1 or x["bonds"][0][0] < 0)):
            return None
        c |= (int(bool(x["use_lds"])) << (2 + 3 * si)) | (int(x["bonds"][0][0] >= 0) << (3 + 3 * si)) | \
             (int(x["bonds"][1][0] >= 0) << (4 + 3 * si))
    return st[-1]["zpair"], c


@pytest.mark.parametrize("L,closed,gx", [(20, True, 0), (20, False, 0), (24, True, 0), (26, True, 0), (20, False, 4)])
def test_sweeps_of_the_benchmarked_chains_have_compiled_kernels(M, L, closed, gx):
    # every sweep of these chains and of the 4 x 5 grid (every layer of the
    # period, every plan variant) and every fused pair (a layer's last sweep,
    # form 0, continuing into the next layer's first in the same window; odd
    # layers run their windows in reverse) has a kernel compiled for its
    # stage structure; a plan without one would fall back to the slower
    # run-time stage loop with no other symptom
    compiled = set(M.compiled_sweeps())
    assert all(zp in (0, 1, 2, 4, 5, 6) for zp, _, _ in compiled)
    period = 4 if gx else 2
    for plan in range(4):
        P = [M.layer_plan(L, L, closed, gx, layer, 13, 0, plan)["passes"] for layer in range(period)]
        for layer in range(period):
            for ps in P[layer]:
                z, c = structure_code(ps)
                assert (z, c, 0) in compiled, (plan, layer, ps["s"], z, c)
            n = len(P[layer])
            last = P[layer][0 if layer & 1 else n - 1]
            nxt = P[(layer + 1) % period]
            first = nxt[len(nxt) - 1 if (layer + 1) & 1 else 0]
            (z0, c0), (z1, c1) = structure_code(last), structure_code(first)
            if last["s"] == first["s"] and z0 == 0 and first["stages"][0]["use_lds"]:
                assert (4 + z1, c0, c1) in compiled, (plan, layer, c0, z1, c1)
            else:
                assert gx, (plan, layer)  # chains always fuse


def test_l20_structure_codes_are_the_three_known_ones(M):
    # the codes of the metric's sweeps, written out: three stages each with
    # bond[0]; form 0 starts in the load mapping, bond[1] in stages 0-1
    # (forms 0, 1) or 0 and 2 (form 2)
    codes = {structure_code(ps) for layer in (0, 1) for ps in M.layer_plan(20, 20, True, 0, layer)["passes"]}
    assert codes == {(0, 3 | (0b110 << 2) | (0b111 << 5) | (0b011 << 8)),
                     (1, 3 | (0b111 << 2) | (0b111 << 5) | (0b011 << 8)),
                     (2, 3 | (0b111 << 2) | (0b011 << 5) | (0b111 << 8))}
