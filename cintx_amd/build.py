"""Build libcintx_hip.so (gfx950) in-tree.

Generates the class-kernel instantiation units, compiles them with hipcc in
parallel and links the shared library next to this file.  Run as
``python -m cintx_amd.build`` or through ``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
GEN = os.path.join(BUILD, "gen")
LIB = os.path.join(HERE, "libcintx_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", CSRC,
]


def class_sets():
    """(family_deriv, form, (li, lj, lk, ll)) instantiated as templates; form
    0 = FP64, 1 = FP32 recurrences, 2 = FP64 with Cartesian output (the
    int2e_cart / int3c2e_cart families)."""
    out = []
    pairs = [(a, b) for a in range(3) for b in range(3) if a >= b]
    # int2e_sph: canonical li>=lj, lk>=ll, (li,lj)>=(lk,ll), l<=2; FP64 and FP32
    for bra in pairs:
        for ket in pairs:
            if bra >= ket:
                for prec in (0, 1):
                    out.append((0, prec, bra + ket))
    # int2e_ip1_sph: derivative on i, only k<->l canonicalised
    for li in range(3):
        for lj in range(3):
            for ket in pairs:
                out.append((1, 0, (li, lj) + ket))
    # int3c2e_sph: li>=lj <= 2, aux l <= 4 in the k slot, FP64 and FP32
    for bra in pairs:
        for lk in range(5):
            for prec in (0, 1):
                out.append((0, prec, bra + (lk, 0)))
    # int3c2e_ip1_sph
    for li in range(3):
        for lj in range(3):
            for lk in range(5):
                out.append((1, 0, (li, lj, lk, 0)))
    # 3c2e (ij|k) with LL == 0 coincides with int2e classes (ij|k s); dedupe
    seen = set()
    uniq = []
    for c in out:
        if c not in seen:
            seen.add(c)
            uniq.append(c)
    uniq += [(0, 2, ls) for deriv, prec, ls in uniq if deriv == 0 and prec == 0]
    return uniq + [(0, 0, ls) for ls in high_l_classes()] + \
        [(1, 0, ls) for ls in high_l_nabla_classes()]


HIGH_L_MAX_NF = 3500
# the wide path builds one chain (G4 values) in registers; up to 135 (with
# the i-row shapes below) every f/g variant compiles without spills
HIGH_L_MAX_G4 = 135


def high_l_classes():
    """int2e_sph classes with an f or g shell (canonical: in the bra), at
    most HIGH_L_MAX_NF Cartesian elements and HIGH_L_MAX_G4 values per
    recurrence chain, FP64 only, without Cartesian or
    fused-J/K copies.  Their blocks are small, so the runtime-class kernel
    (one tuple per 256-lane workgroup, one primitive tuple per step) leaves
    most of its lanes idle on them; the template kernels batch tuples and
    primitive tuples instead.  A bra whose f or g shell pairs with a shell
    of l >= 2 runs the wide path with the transfer base chosen per
    primitive tuple (needs_pb); libcint's fixed base loses up to 1e-10 in
    FP64 there (144x the 1e-12 bar on (g g|s s) tuples of the s..g
    molecule, 12x on (g g|p s)).  Kets stay s..d or an f/g shell with an s
    or p partner (no per-primitive ket base); larger classes keep the
    runtime-class kernel."""
    nc = [(l + 1) * (l + 2) // 2 for l in range(5)]
    pairs = [(a, b) for a in range(5) for b in range(5) if a >= b]
    out = []
    for bra in pairs:
        for ket in pairs:
            ls = bra + ket
            g4 = (ls[0] + 1) * (ls[1] + 1) * (ls[2] + 1) * (ls[3] + 1)
            if bra >= ket and max(ls) >= 3 and (ket[0] < 3 or ket[1] < 2) and \
                    nc[ls[0]] * nc[ls[1]] * nc[ls[2]] * nc[ls[3]] <= HIGH_L_MAX_NF and \
                    g4 <= HIGH_L_MAX_G4:
                out.append(ls)
    return out


def high_l_nabla_classes():
    """nabla_i classes with an f or g shell i and an s shell j, kets s..d:
    int3c2e_ip2_sph runs as nabla_i on (k | i j) with the aux shell k in the
    i slot (and an empty j), and int2e_ip1_sph has them for (f s|..) and
    (g s|..).  With j = s the bra needs no transfer; libcint's FP64 route
    stays within a sixtieth (int3c2e_ip2) to a half (int2e_ip1 (f s|p s))
    of the 1e-12 bar on tuples of neighbouring atoms."""
    pairs = [(a, b) for a in range(3) for b in range(3) if a >= b]
    return [(li, 0) + ket for li in (3, 4) for ket in pairs]


def needs_pb(ls):
    """f or g bra shell paired with l >= 2: per-primitive transfer base"""
    return max(ls[0], ls[1]) >= 3 and min(ls[0], ls[1]) >= 2


def is_high_l(deriv, ls):
    return max(ls[0], ls[1]) >= 3


def _p2(v):
    p = 1
    while p < v:
        p *= 2
    return p


def register_chunks(deriv, prec, ls):
    """Row chunks for the one-lane path on a class too large for one pass:
    the g tensor of one root (3*G4 values) and a chunk's FP64 accumulators
    should fit about 130 doubles of registers (VGPRs plus AGPRs at one wave
    per SIMD; measured: beyond that the compiler spills tens of KB to scratch
    and the kernel runs 100x slower).  Every chunk is an unrolled copy of the primitive loop, so at
    most 3 (more thrash the instruction cache: 6 chunks ran 100x slower);
    None when that is not enough."""
    li, lj, lk, ll = ls
    nc = [(l + 1) * (l + 2) // 2 for l in ls]
    g4 = (li + deriv + 1) * (lj + 1) * (lk + 1) * (ll + 1)
    gdoubles = 3 * g4 * (8 if prec == 0 else 4) / 8
    nb = nc[0] if deriv else nc[0] * nc[1]
    nrow = nc[0] * nc[1] * nc[2] * nc[3] // nb
    if gdoubles > 72:
        return None
    # FP32 keeps a per-primitive-tuple partial sum beside each FP64 accumulator
    per_acc = 1.0 if prec == 0 else 1.5
    for nch in (2, 3):
        rpc = -(-nrow // nch)
        if gdoubles + per_acc * rpc * nb * (3 if deriv else 1) <= 130:
            return nch
    return None


def variants(deriv, prec, ls):
    """Launch-shape candidates (lanes per tuple T, primitive tuples per step P)
    of a class kernel; the first is the untuned default.  Mirrors the sizes of
    ClassShape in csrc/eri_device.h."""
    if prec == 2:  # Cartesian output: the FP64 shapes with PP bit 11
        return [(tt, pp | 2048) for tt, pp in variants(deriv, 0, ls)]
    li, lj, lk, ll = ls
    nc = [(l + 1) * (l + 2) // 2 for l in ls]
    nr = (li + deriv + lj + lk + ll) // 2 + 1
    g4 = (li + deriv + 1) * (lj + 1) * (lk + 1) * (ll + 1)
    gs = 3 * g4 * nr
    nf = nc[0] * nc[1] * nc[2] * nc[3]
    ne = nf * (3 if deriv else 1)
    if ne <= 8:
        return [(1, 1)]
    nb = nc[0] if deriv else nc[0] * nc[1]
    nrow = nf // nb
    narrow = 4 if ne <= 48 else 16 if ne <= 192 else 64 if ne <= 768 else 256
    wide = min(256, max(4, _p2(max(nrow, 3 * nr))))
    groups = wide // nrow if wide >= nrow else 1
    plds = max(1, (60 * 1024) // (gs * 8 + (3 * nr + 16) * 8))
    p0 = max(1, min(8, wide // (3 * nr)))
    pw = min(plds, max(p0, min(groups, 8)))
    cands = [(wide, pw), (wide, min(plds, 2 * pw, 8)), (wide, 1), (narrow, 1)]
    # twice and half the lanes (more primitive slots in flight, or more rows
    # per lane)
    w2 = min(256, 2 * wide)
    plds2 = max(1, (60 * 1024) // (gs * 8 + (3 * nr + 16) * 8))
    cands.append((w2, max(1, min(8, w2 // (3 * nr), plds2))))
    if wide >= 8:
        cands.append((wide // 2, 1))
    if not deriv and ne > 81:
        # rows of the i components alone (P | 256): more lanes per tuple
        nrow_i = nf // nc[0]
        ti = min(256, max(4, _p2(max(nrow_i, 3 * nr))))
        pi = max(1, min(8, ti // (3 * nr), plds))
        cands.insert(2, (ti, pi | 256))
    # one lane per tuple, everything in registers: the default for small
    # classes, a candidate up to 81 elements
    # (a floor of 2-3 waves per SIMD, PP bits 9-10, was measured: the
    # register kernels then spill and run 1-7x slower, so it is not offered)
    if ne <= 27:
        cands.insert(0, (1, 1))
    elif ne <= 81:
        cands.insert(3, (1, 1))
    elif not is_high_l(deriv, ls):
        # (f/g classes: no chunked one-lane variant; with the promoted-alloca
        # limits below their compile time runs to tens of minutes)
        nch = register_chunks(deriv, prec, ls)
        if nch:
            cands.insert(3, (1, nch))
    if is_high_l(deriv, ls) and not deriv and nb > 36:
        # f/g bras with more (i, j) components than (d d|: rows of the i
        # components alone (PP bit 8), so a lane's accumulators stay in
        # registers (rows of 60-225 elements spilled 1.7-2.9 KB per lane and
        # ran up to 50x slower)
        nrow_i = nf // nc[0]
        ti = min(256, max(4, _p2(max(nrow_i, 3 * nr))))
        t2 = min(256, 2 * ti)
        cands = [(ti, max(1, min(8, ti // (3 * nr), plds)) | 256), (ti, 1 | 256),
                 (t2, max(1, min(8, t2 // (3 * nr), plds)) | 256), (max(4, ti // 2), 1 | 256)]
    if is_high_l(deriv, ls) and needs_pb(ls):
        # wide path only, with the per-primitive bra base (PP bit 13)
        cands = [(tt, pp | 8192) for tt, pp in cands if tt > 1]
    out = []
    for c in cands:
        if c not in out:
            out.append(c)
    if ne <= 27:  # the register path won every such class by 3-10x
        return out[:2]
    return out[:6]


# Translation units of the chunked one-lane variants are compiled with larger
# alloca-to-register limits: their g tensor (up to 81 doubles) is indexed with
# compile-time constants only after full unrolling, and the default limits
# leave it in scratch memory.
CHUNK_FLAGS = ["-mllvm", "-amdgpu-promote-alloca-to-vector-limit=4096",
               "-mllvm", "-amdgpu-promote-alloca-to-vector-max-regs=512",
               "-mllvm", "-amdgpu-promote-alloca-to-vector-vgpr-ratio=1"]


def generate(nparts=24, nchunkparts=12):
    os.makedirs(GEN, exist_ok=True)
    classes = class_sets()
    plain, chunked = [], []
    for deriv, prec, ls in classes:
        for v, (tt, pp) in enumerate(variants(deriv, prec, ls)):
            (chunked if (tt == 1 and (pp & 255) > 1) else plain).append((deriv, prec, ls, v, tt, pp))
            # the same launch shape with the fused J/K epilogue (PP bit 12):
            # int2e_sph classes in FP64 and FP32 (not the f/g classes)
            if deriv == 0 and prec in (0, 1) and not is_high_l(deriv, ls):
                (chunked if (tt == 1 and (pp & 255) > 1) else plain).append(
                    (deriv, prec, ls, v, tt, pp | 4096))
    parts = [plain[i::nparts] for i in range(nparts)]
    parts += [chunked[i::nchunkparts] for i in range(nchunkparts)]
    nparts = len(parts)
    files = []
    for p, cls in enumerate(parts):
        lines = ['#include "eri_device.h"', "namespace cx {",
                 f"void register_part_{p}(Registry& reg) {{"]
        for deriv, prec, (a, b, c, d), v, tt, pp in cls:
            real = "float" if prec == 1 else "double"
            slot = (f"jk[{prec}]" if pp & 4096 else f"k[{deriv}][{prec}]") + f"[{a}][{b}][{c}][{d}][{v}]"
            lines.append(f"  reg.{slot} = "
                         f"class_info<{a}, {b}, {c}, {d}, {deriv}, {real}, {tt}, {pp}>();")
        lines += ["}", "}  // namespace cx", ""]
        stem = f"inst_{p:02d}" if p < len(parts) - nchunkparts else f"inst_reg_{p:02d}"
        path = os.path.join(GEN, stem + ".hip")
        text = "\n".join(lines)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)
        files.append(path)
    reg = ['#include "launch.h"', "namespace cx {"]
    reg += [f"void register_part_{p}(Registry& reg);" for p in range(nparts)]
    reg += ["void register_class_kernels(Registry& reg) {"]
    reg += [f"  register_part_{p}(reg);" for p in range(nparts)]
    reg += ["}", "}  // namespace cx", ""]
    path = os.path.join(GEN, "inst_registry.cpp")
    text = "\n".join(reg)
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    files.append(path)
    return files


def _compile(src, obj, extra):
    deps = [src, os.path.join(CSRC, "eri_device.h"), os.path.join(CSRC, "common.h"),
            os.path.join(CSRC, "jk_device.h"),
            os.path.join(CSRC, "launch.h"), os.path.join(CSRC, "rys_table.h"),
            os.path.join(HERE, "..", "include", "cintx_hip.h")]
    if os.path.exists(obj):
        t = os.path.getmtime(obj)
        if all(os.path.getmtime(d) <= t for d in deps if os.path.exists(d)):
            return obj, None
    cmd = [HIPCC] + COMMON_FLAGS + extra + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON_FLAGS + extra + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, (" ".join(cmd)) + "\n" + r.stdout + r.stderr
    return obj, None


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    gen = generate()
    srcs = [os.path.join(CSRC, "runtime.cpp"), os.path.join(CSRC, "rys_table.cpp"),
            os.path.join(CSRC, "rys_kernels.hip"), os.path.join(CSRC, "eri_generic.hip"),
            os.path.join(CSRC, "rys_mlp.hip"), os.path.join(CSRC, "jk_digest.hip"),
            os.path.join(CSRC, "spinor.hip")] + gen
    extra = {os.path.join(CSRC, "rys_table.cpp"): ["-fopenmp"]}
    for g in gen:
        if os.path.basename(g).startswith("inst_reg_"):
            extra[g] = CHUNK_FLAGS
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs = []
    errors = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = []
        for s in srcs:
            obj = os.path.join(BUILD, os.path.basename(s) + ".o")
            futs.append(ex.submit(_compile, s, obj, extra.get(s, [])))
        for f in futs:
            obj, err = f.result()
            objs.append(obj)
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        nk = sum(len(variants(d, p, l)) for d, p, l in class_sets())
        print(f"built {LIB} ({len(class_sets())} classes, {nk} class-kernel variants)")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
