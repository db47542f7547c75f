"""Helpers shared by the GPU parity tests: run the HIP engine and the CPU oracle
side by side and compare their complete states bitwise."""
import numpy as np

import oracle as O

# GPU field name -> oracle field index
FIELD_MAP = {"U": O.U, "Us": O.US, "Bi": O.BI, "Bis": O.BIS, "ff": O.FF, "gg": O.GG, "EMF": O.EMF}


def oracle_state(d):
    lp, ic, lf = d.get_tree()
    st = dict(lp=lp, icoord=ic, leafs=lf)
    for name, f in FIELD_MAP.items():
        st[name] = d.get_field(f)
    inf = d.info()
    st["t"], st["dt"] = inf["t"], inf["dt"]
    return st


def compare_states(gpu_st, ora_st, fields=tuple(FIELD_MAP)):
    """Return a list of human-readable differences (empty when bitwise equal)."""
    diffs = []
    for k in ("lp", "icoord", "leafs"):
        if gpu_st[k].shape != ora_st[k].shape or not np.array_equal(gpu_st[k], ora_st[k]):
            diffs.append(f"tree array {k} differs")
    for f in fields:
        a, b = gpu_st[f], ora_st[f]
        if a.shape != b.shape:
            diffs.append(f"{f}: shape {a.shape} vs {b.shape}")
            continue
        ne = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if ne.any():
            idx = np.argwhere(ne)
            nb, k, j, i = idx[0]
            diffs.append(f"{f}: {int(ne.sum())} elements differ, first at slot {nb} var {k} (i={i}, j={j}): "
                         f"gpu {a[nb, k, j, i]!r} oracle {b[nb, k, j, i]!r}")
    return diffs


def gpu_state(sim, inst=0):
    st = sim.state(inst)
    inf = sim.info(inst)
    st["t"], st["dt"] = inf["t"], inf["dt"]
    return st


def lockstep(sim, dom, steps, full_every=10, label=""):
    """Advance GPU and oracle together; compare tree/time every step and all
    fields every `full_every` steps and at the end.  Returns steps done."""
    done = 0
    for s in range(steps):
        if dom.done():
            break
        rc_o = dom.step()
        assert rc_o == 0, f"{label}: oracle error {rc_o} {dom.last_error()} at step {s}"
        sim.step()
        done += 1
        gi, oi = sim.info(), dom.info()
        assert gi["last_active"] == oi["lastActive"] and gi["nleaf"] == oi["nleaf"], (
            f"{label}: tree size differs at step {s}: gpu {gi} oracle {oi}")
        assert gi["t"] == oi["t"] and gi["dt"] == oi["dt"] and gi["dt_next"] == oi["dt_next"], (
            f"{label}: time differs at step {s}: gpu {gi} oracle {oi}")
        if (s + 1) % full_every == 0 or dom.done() or s == steps - 1:
            d = compare_states(gpu_state(sim), oracle_state(dom))
            assert not d, f"{label}: state differs after step {s}: " + "; ".join(d[:4])
    return done


def hashes_via_oracle(sim, dom, fields=True):
    """FNV-1a hashes of the GPU's quadtree and (with `fields`) of every slot
    field, computed by the oracle's hash functions over the downloaded state
    (dom: an oracle Domain of the same configuration, used for hashing only)."""
    lp, ic, lf = sim.get_tree()
    dom.set_tree(lp, ic, lf)
    if not fields:
        return None, dom.hash_tree()
    for name, f in FIELD_MAP.items():
        dom.set_field(f, sim.download(name))
    return dom.hash_fields(), dom.hash_tree()


def follow_oracle_log(sim, dom, rows, label):
    """Step the GPU through an oracle-written log (tests/golden/
    make_oracle_logs.py): per step the applied dt, t, slot and leaf counts and
    the tree hash; the field hash wherever the log holds one."""
    r = rows[0]
    hf, ht = hashes_via_oracle(sim, dom, fields=bool(r["hash_fields"]))
    assert ht == int(r["hash_tree"]), f"{label}: initial tree differs"
    if r["hash_fields"]:
        assert hf == int(r["hash_fields"]), f"{label}: initial fields differ"
    for r in rows[1:]:
        applied = sim.info()["dt"]
        sim.step()
        inf = sim.info()
        got = (applied, inf["t"], inf["last_active"], inf["nleaf"])
        want = (r["dt"], r["t"], r["last_active"], r["nleaf"])
        assert got == want, f"{label}: step {int(r['step'])}: gpu {got} oracle {want}"
        hf, ht = hashes_via_oracle(sim, dom, fields=bool(r["hash_fields"]))
        assert ht == int(r["hash_tree"]), f"{label}: tree differs after step {int(r['step'])}"
        if r["hash_fields"]:
            assert hf == int(r["hash_fields"]), f"{label}: fields differ after step {int(r['step'])}"
