"""Extract the known answers of three more reference unit suites into JSON (data only).

Run in a container that holds the reference sources:
    python tests/golden/extract_reference_unit_kat.py /root/reference \
        tests/golden/reference_unit_kat.json

It reads the Catch2 sources as text and keeps the values they hold and the answers they expect:
  * test/units/simulation_internal/extrusion_barriers_test.cpp: the (stp_inactive, occupancy)
    pairs of "occupancy", the (stp_active, stp_inactive, expected, approximate) rows of
    "compute_occupancy" and the positions of "sort";
  * test/units/contact_matrix/contact_matrix_dense_test.cpp: the matrix shape and the sequence of
    increments, decrements and checks of "CMatrix simple" and "CMatrix in/decrement" (the checks
    under utils::ndebug_not_defined(), which expect exceptions of debug builds, are left out), and
    the fills and unsafe_get_block checks of "CMatrix get w/ block";
  * test/units/simulation_cpu/collision_encoding_test.cpp: every (index, event) the cases encode,
    each collision_occurred / collision_avoided / decode_* check with its expected answer, in
    order, as a list of operations on one collision word (the checks of invalid events under
    utils::ndebug_not_defined() are left out).
Nothing from the reference is executed.
"""
import json
import re
import sys

EV = {"COLLISION": 0x10, "CHROM_BOUNDARY": 0x08, "LEF_BAR": 0x04, "LEF_LEF_PRIMARY": 0x02,
      "LEF_LEF_SECONDARY": 0x01}
U32F_BITS = 64  # u32f is uint_fast32_t: 64 bits on x86-64 Linux (the reference's target)


def cases_of(text):
    out = {}
    for m in re.finditer(r'TEST_CASE\("([^"]+)"', text):
        i = text.index("{", m.end())
        depth, j = 0, i
        while True:
            depth += {"{": 1, "}": -1}.get(text[j], 0)
            if depth == 0:
                break
            j += 1
        out[m.group(1)] = text[i + 1:j]
    return out


def strip(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def drop_debug_blocks(body):
    """Removes `if constexpr (utils::ndebug_not_defined()) { ... }` blocks."""
    while True:
        m = re.search(r"if constexpr \(utils::ndebug_not_defined\(\)\)\s*\{", body)
        if not m:
            return body
        depth, j = 0, m.end() - 1
        while True:
            depth += {"{": 1, "}": -1}.get(body[j], 0)
            if depth == 0:
                break
            j += 1
        body = body[:m.start()] + body[j + 1:]


def event_of(expr):
    v = 0
    for name in re.findall(r"Collision<>::(\w+)", expr):
        v |= EV[name]
    return v


def barriers(text):
    cs = cases_of(strip(text))
    occ = cs["Extrusion barriers - occupancy"]
    vals = dict(re.findall(r"const auto (\w+) = ([0-9.]+);", occ))
    pairs = [{"stp_inactive": float(vals[f"stp_inactive{k}"]), "occupancy": float(vals[f"occupancy{k}"])}
             for k in (1, 2)]
    rows = []
    body = cs["Extrusion barriers - compute_occupancy"]
    stp_a = stp_i = None
    for stmt in body.split(";"):
        m = re.search(r"stp_active = ([0-9.]+)\s*$", stmt)
        if m:
            stp_a = float(m.group(1))
        m = re.search(r"stp_inactive = ([0-9.]+)\s*$", stmt)
        if m:
            stp_i = float(m.group(1))
        m = re.search(r"compute_occupancy_from_stp\(stp_active, stp_inactive\)\s*==\s*"
                      r"(Catch::Approx\()?([0-9.]+)", stmt)
        if m:
            rows.append({"stp_active": stp_a, "stp_inactive": stp_i, "expected": float(m.group(2)),
                         "approx": bool(m.group(1))})
    m = re.search(r"ExtrusionBarriers\{\{([0-9, ]+)\}", cs["Extrusion barriers - sort"])
    positions = [int(x) for x in m.group(1).split(",")]
    return {"occupancy": pairs, "compute_occupancy": rows, "sort_positions": positions}


def cmatrix(text):
    cs = cases_of(strip(text))
    out = {}
    for name in ("CMatrix simple", "CMatrix in/decrement"):
        body = drop_debug_blocks(cs[name])
        m = re.search(r"ContactMatrixDense<\w*>\s+(\w+)\((\d+),\s*(\d+)\)", body)
        var, nrows, ncols = m.group(1), int(m.group(2)), int(m.group(3))
        ops = []
        for stmt in body.split(";"):
            stmt = stmt.strip()
            if m := re.fullmatch(rf"{var}\.(increment|decrement)\((\d+),\s*(\d+)\)", stmt):
                ops.append([m.group(1), int(m.group(2)), int(m.group(3)), 1])
            elif m := re.fullmatch(rf"{var}\.subtract\((\d+),\s*(\d+),\s*(\d+)\)", stmt):
                ops.append(["decrement", int(m.group(1)), int(m.group(2)), int(m.group(3))])
            elif m := re.fullmatch(rf"(?:CHECK|REQUIRE)\({var}\.get\((\d+),\s*(\d+)\) == (\d+)\)", stmt):
                ops.append(["get", int(m.group(1)), int(m.group(2)), int(m.group(3))])
            elif m := re.fullmatch(rf"(?:CHECK|REQUIRE)\({var}\.get_tot_contacts\(\) == (\d+)\)", stmt):
                ops.append(["tot", int(m.group(1))])
            elif m := re.fullmatch(rf"(?:CHECK|REQUIRE)\({var}\.get_n_of_missed_updates\(\) == (\d+)\)",
                                   stmt):
                ops.append(["missed", int(m.group(1))])
            elif stmt and not stmt.startswith("ContactMatrixDense"):
                raise ValueError(f"{name}: unhandled statement {stmt!r}")
        out[name] = {"nrows": nrows, "ncols": ncols, "ops": ops}
    out["CMatrix get w/ block"] = cmatrix_block(cs["CMatrix get w/ block"])
    return out


def cmatrix_block(body):
    """The fills (nested loops over i and j with the value set at (i, j), kept as the loop bounds
    and the value's arithmetic expression in i and j) and the unsafe_get_block checks."""
    m = re.search(r"ContactMatrixDense<\w*>\s+(\w+)\((\d+),\s*(\d+)\)", body)
    var, nrows, ncols = m.group(1), int(m.group(2)), int(m.group(3))
    fills = []
    loop = (r"for \(u32 i = (\d+); i < (\d+); \+\+i\)\s*\{\s*"
            r"for \(u32 j = (\w+); j < (\d+); \+\+j\)\s*\{\s*"
            rf"{var}\.set\(i, j, ([^;]+)\);")
    for m in re.finditer(loop, body):
        j0 = m.group(3)
        fills.append({"i": [int(m.group(1)), int(m.group(2))],
                      "j": [j0 if j0 == "i" else int(j0), int(m.group(4))],
                      "value": " ".join(m.group(5).split())})
    checks = [[int(a), int(b), int(c), int(d)] for a, b, c, d in re.findall(
        rf"CHECK\({var}\.unsafe_get_block\((\d+), (\d+), (\d+)\) == (\d+)\)", body)]
    if not fills or not checks:
        raise ValueError("CMatrix get w/ block: nothing extracted")
    return {"nrows": nrows, "ncols": ncols, "fills": fills, "block_checks": checks}


def collisions(text):
    text = strip(text)
    bits = int(re.search(r"RESERVED_EVENT_BITS = (\d+);", text).group(1))
    index_bits = U32F_BITS - bits
    index_mask = (1 << (index_bits - 1)) - 1  # ~(~0 << (INDEX_BITS - 1))
    out = {"index_mask": index_mask, "cases": {}}
    for name, body in cases_of(text).items():
        body = drop_debug_blocks(body)
        ops = []
        m = re.search(r"std::array<usize, \d+>\{([0-9, ]+)\}", body)
        loop_idx = [int(x) for x in m.group(1).split(",")] if m else [None]
        for idx0 in loop_idx:
            idx, event = idx0, None
            for stmt in body.split(";"):
                stmt = " ".join(stmt.split())
                if m := re.search(r"(?:const )?usize idx = (\d+)$", stmt):
                    idx = int(m.group(1))
                elif m := re.search(r"^idx = (\w+)$", stmt):
                    idx = index_mask if m.group(1) == "INDEX_MASK" else int(m.group(1))
                elif m := re.search(r"(?:auto|CollisionEvent<u8f>)?\s*\bevent = (.+)$", stmt):
                    if "Collision<>::" in m.group(1):
                        event = event_of(m.group(1))
                elif m := re.search(r"std::array<CollisionEvent<u8f>, \d+> events\{(.*)\}$", stmt, re.S):
                    for ev in m.group(1).split(","):
                        ops.append(["roundtrip", idx, event_of(ev)])
                elif m := re.fullmatch(r"(CHECK|CHECK_FALSE|REQUIRE)\(Collision<>\{idx, event\}\."
                                       r"collision_(occurred|avoided)\((?:Collision<>::(\w+))?\)\)",
                                       stmt):
                    ops.append(["check", idx, event, m.group(2), EV[m.group(3)] if m.group(3) else 0,
                                m.group(1) != "CHECK_FALSE"])
                elif m := re.fullmatch(r"collision\.set_event\((.+)\)", stmt):
                    ops.append(["set_event", event_of(m.group(1))])
                elif m := re.fullmatch(r"collision\.set_idx\(idx\)", stmt):
                    ops.append(["set_idx", idx])
                elif m := re.fullmatch(r"collision\.set\(idx, (.+)\)", stmt):
                    ops.append(["set", idx, event_of(m.group(1))])
                elif m := re.fullmatch(r"(CHECK|REQUIRE)\(collision\.collision_(occurred|avoided)"
                                       r"\(Collision<>::(\w+)\)\)", stmt):
                    ops.append(["word_check", m.group(2), EV[m.group(3)], True])
                elif re.fullmatch(r"CHECK\(collision\.decode_index\(\) == idx\)", stmt):
                    ops.append(["word_index", idx])
        if ops:
            out["cases"][name] = ops
    return out


def main(ref_root, out_path):
    src = {
        "barriers": "test/units/simulation_internal/extrusion_barriers_test.cpp",
        "cmatrix": "test/units/contact_matrix/contact_matrix_dense_test.cpp",
        "collisions": "test/units/simulation_cpu/collision_encoding_test.cpp",
    }
    res = {}
    for key, rel in src.items():
        with open(f"{ref_root}/{rel}") as fh:
            text = fh.read()
        res[key] = {"barriers": barriers, "cmatrix": cmatrix, "collisions": collisions}[key](text)
        res[key]["source"] = rel
    with open(out_path, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")
    print(f"wrote {out_path}")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
