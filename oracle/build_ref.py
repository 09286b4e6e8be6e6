"""Recipe: build the reference Fortran program into oracle/_ref/.

Reads the reference source where it lies under /root/reference, applies the
mechanical patches listed below in memory, compiles the patched text from a
temporary file under /tmp with AMD flang, and writes only the executables into
oracle/_ref/.  No reference source is copied into the repository or into
oracle/_ref/.

Patches (each is a parameter or syntax edit, none changes the algorithm):
  * IF(0)/IF(1) -> IF(.FALSE.)/IF(.TRUE.): flang rejects integer conditions
    (F6:223, F6:596, F6:633, F6:690).
  * F6:1481 is commented out: a constant out-of-bounds subscript inside the
    never-called WRITE_RESULTSPROFILE that flang rejects at compile time.
  * PAUSE statements removed (F6:113, F6:197): the program would block on stdin.
  * PARAMETER values: NX, NY (F6:8-9), NON_NEWTONIAN (F6:4), WRITE_OUTPUT
    (F6:6, so the final state is dumped as list-directed text, which flang
    prints in shortest round-trip form), TIME_MAX (F6:41).
  * Optional: the top-wall velocity UXN (F6:1204) for a moving-lid case.
  * Optional ("local" rate mode): one line at the top of the Collision_MRT node
    loop (after F6:877) that loads the node's own previous-step rate,
    CALL ComputeRelaxation_MRT(OMEGA_BN(I,J)).
  * Optional ("mohammad" rates): in ComputeRelaxation_MRT the active rate
    lines (F6:1009-1011) are commented out and the commented "Mohammad MRT"
    set (F6:1017-1018) is enabled, the rate set of the one-disk Newtonian
    workload (also F1:1044-1045).
  * Optional ("zouhe" boundary): F6's BOUNDARYCONDITION (F6:1195-1390) is
    replaced by F1's (F1:1222-1417: Zou-He edges, bounce-back corners) with
    F6's argument list, and optionally a top-wall velocity in F1:1230.
  * Optional ("equilibrium" boundary): the same with F4's routine
    (F4:1205-1402: every edge and corner node set to the equilibrium of its
    interior neighbour); its first two comment-only lines (F4:1216-1217)
    are dropped so the routine keeps F6's length.
  * Optional ("casson" rheology): in Collision_MRT the two Papanastasiou-
    Bingham viscosity lines (F6:940, F6:943) are commented out and the two
    Casson lines F6 carries commented next to them (F6:941, F6:944; the same
    model as F3:986-989) are enabled.

A second recipe builds the N-body contact force alone: MODULE PARAM
(F3:2-83) and SUBROUTINE ComputeCollisionForce (F3:632-806) cut from F3,
compiled with oracle/f3_contact_driver.f90 into oracle/_ref/f3_contact.
A further recipe builds the Wan-Turek and Lennard-Jones contact models:
MODULE PARAM and ComputeCollisionForce(DISK,WALL) cut from F4 (F4:2-89,
F4:576-743) or F1 (F1:2-93, F1:596-769), the routine's model switches
(F4:587, F4:624, F4:680; F1:609, F1:646, F1:701) set to the model under test,
compiled with oracle/contact_model_driver.f90 into oracle/_ref/f4_contact_wt
(F4 as it ships: Wan-Turek, zeta 1), f4_contact_lj and f1_contact_wt (zeta 2).
A fourth recipe builds whole multi-particle programs (patch_nparticle): F3
(grid, particle count and steps set, the STOP before its time loop at
F3:202 commented out, two output lines added) into oracle/_ref/f3_multi,
which reads its disks from 0_Particle_Initial_Position.txt in the run
directory, and F2, the two-disk program (grid, steps and its two initial
centres set, the same output lines) into oracle/_ref/f2_dkt.  Both are
compiled with -fno-automatic (see below).
A third recipe builds the reference's own MAIN program on the GPU: F6
(patched as its variant) with the nine time-step subroutines that
2d_nnf_lbm_amd/fortran/lbm_f6_shim.F90 replaces cut out, compiled with that
shim and the LBM_AMD module and linked to liblbm_amd.so, into
oracle/_ref/f6_shim_<variant>.  The MAIN program and every other routine
are the reference's, unchanged.

It is compiled with -fno-automatic, which gives the routine's local variables
the SAVE attribute: the pair branch (F3:760) reads D_PW before this call
assigns it, so it sees the value the previous call left (the "stale D_PW"
the oracle and the GPU path implement).  Without SAVE the value is whatever
the stack held, i.e. undefined.

Usage: python oracle/build_ref.py [--timing]   (no-op when /root/reference is
absent, e.g. on a machine that only runs the prebuilt files).
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

REF_ROOT = "/root/reference"
F6 = os.path.join(REF_ROOT, "6_FreeFall_YieldStress",
                  "D2Q9_BinghamSediment_Validated_20191127.f90")
F1 = os.path.join(REF_ROOT, "1_Disk_Newtonian", "D2Q9_1P_Validated_20190530.f90")
F3 = os.path.join(REF_ROOT, "3_MultiParticle_Sedimentation",
                  "D2Q9_RTParticle_20191003.f90")
F2 = os.path.join(REF_ROOT, "2_DKT_Newtonian", "D2Q9_DKT_Validated_20190926.f90")
F4 = os.path.join(REF_ROOT, "4_Oscillating_Particle",
                  "D2Q9_V7_OscillatingCylinder_Validated.f90")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref")
FLANG = shutil.which("flang") or "/opt/rocm/llvm/bin/flang"

# name -> patch settings.  Parity variants are small grids run to 1000 steps.
VARIANTS = {
    "f6_newt": dict(nx=40, ny=120, non_newtonian=False, time_max=1000),
    "f6_bing_lagged": dict(nx=40, ny=120, non_newtonian=True, time_max=1000),
    "f6_bing_local": dict(nx=40, ny=120, non_newtonian=True, time_max=1000,
                          local_rate=True),
    "f6_newt_lid": dict(nx=40, ny=60, non_newtonian=False, time_max=300,
                        uxn="0.01D0"),
    "f6_newt_mohammad": dict(nx=40, ny=120, non_newtonian=False, time_max=1000,
                             rates="mohammad"),
    "f6_cas_local": dict(nx=40, ny=120, non_newtonian=True, time_max=1000,
                         local_rate=True, rheology="casson"),
    "f6_zouhe_lid": dict(nx=40, ny=60, non_newtonian=False, time_max=300,
                         boundary="zouhe", zouhe_uxn="0.01D0"),
    "f6_eq_bing_local": dict(nx=40, ny=120, non_newtonian=True, time_max=1000,
                             local_rate=True, boundary="equilibrium"),
}
# CPU-baseline builds: the reference's Bingham settling disk on the 401x1601
# node grid (NX=400, NY=1600), timed for a bounded number of steps; its
# Newtonian single disk on the one-disk workload's 201x601 grid with the
# alternative rate set.
TIMING_VARIANTS = {
    "f6_bing_400x1600": dict(nx=400, ny=1600, non_newtonian=True, time_max=40,
                             write_output=False),
    "f6_newt_200x600": dict(nx=200, ny=600, non_newtonian=False, time_max=600,
                            write_output=False, rates="mohammad"),
}
# The oscillating-cylinder program (F4) on square cavities of three sides
# (the ensemble's size range), Papanastasiou fluid, output files off, the
# time loop bounded; its characteristic length L_C (cylinder diameter, and
# the inner wall box of 10 x 15 L_C, F4:179-180) set to side/30 so the box
# lies inside the cavity.  name -> (side, steps)
F4_TIMING = {"f4_cav_128": (128, 2000), "f4_cav_512": (512, 100), "f4_cav_1024": (1024, 25)}
# The multi-particle program (F3) with the dense workload's 504 disks on the
# 2048 x 2048 cell cavity, three steps (its per-particle full-grid solid
# fractions take about 17 GB).
F3_TIMING = {"f3_dense_2048": dict(src="F3", nx=2048, ny=2048, n_particle=504, time_max=3,
                                   clock=True, write_output=False)}


def patch_f4_timing(text, side, steps):
    """F4 as shipped with NX = NY = side, NON_NEWTONIAN and WRITE_OUTPUT
    switched (F4:4, F4:6), L_C = side/30 (F4:27), the time loop bounded to
    `steps` iterations (F4:216) and IF(0)/IF(1) written as logicals."""
    reps = [
        (r"(INTEGER,PARAMETER :: NX=)\d+", r"\g<1>%d" % side),
        (r"(INTEGER,PARAMETER :: NY=)\d+", r"\g<1>%d" % side),
        (r"(LOGICAL,PARAMETER :: NON_NEWTONIAN=)\s*\.FALSE\.", r"\g<1> .TRUE."),
        (r"(LOGICAL,PARAMETER :: WRITE_OUTPUT=)\s*\.TRUE\.", r"\g<1> .FALSE."),
        (r"(DOUBLE PRECISION,PARAMETER :: L_C=)50\.0", r"\g<1>%.1f" % (side / 30.0)),
        (r"DO WHILE\(KK<INT\(6\.0\*T_C\)\)", "DO WHILE(KK<%d)" % steps),
    ]
    for pat, rep in reps:
        text, n = re.subn(pat, rep, text, count=1)
        if n != 1:
            raise RuntimeError("F4 timing patch did not apply: " + pat)
    # integer IF conditions as logicals (flang rejects the integer form)
    text = re.sub(r"\bIF\(0\)", "IF(.FALSE.)", text)
    text = re.sub(r"\bIF\(1\)", "IF(.TRUE.)", text)
    # the program's CPU clock (START is set at F4:135) written after every
    # loop iteration, so a timing can leave the set-up out
    text, n = re.subn(r"(\n(\s*)WRITE\(8,\*\) KK,AvgDen[^\n]*)",
                      r"\1\n\2CALL CPU_TIME(FINISH);WRITE(98,*) KK,FINISH-START;FLUSH(98)",
                      text, count=1)
    if n != 1:
        raise RuntimeError("F4 timing patch: the loop's density line was not found")
    return text


def build_f4_timing(name, side, steps, flags):
    with open(F4, "rb") as fh:
        text = fh.read().decode("ascii", errors="replace").replace("\r", "")
    src = patch_f4_timing(text, side, steps)
    os.makedirs(OUT_DIR, exist_ok=True)
    tmpdir = tempfile.mkdtemp(prefix="refbuild_", dir="/tmp")
    try:
        path = os.path.join(tmpdir, name + ".f90")
        with open(path, "w") as fh:
            fh.write(src)
        exe = os.path.join(OUT_DIR, name)
        r = subprocess.run([FLANG] + flags + [path, "-o", exe], cwd=tmpdir,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode:
            raise RuntimeError("%s build failed:\n%s" % (name, r.stdout[-3000:]))
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return exe


def patch_source(text, nx, ny, non_newtonian, time_max, local_rate=False,
                 uxn=None, write_output=True, rates=None, rheology=None):
    lines = text.replace("\r", "").split("\n")
    out = []
    in_collide = False
    in_rate = False
    in_relax = False
    enable = 0
    for n, line in enumerate(lines, start=1):
        s = line
        if rates == "mohammad":
            if re.match(r"\s*SUBROUTINE ComputeRelaxation_MRT\(", s):
                in_relax = True
            elif in_relax and re.match(r"\s*END SUBROUTINE ComputeRelaxation_MRT", s):
                in_relax = False
            elif in_relax:
                if re.match(r"\s*(omega_e=1\.1;omega_eps=1\.25|Omega_mrt\(0\)=0\.0;|Omega_mrt\(4\)=omega_q;)", s):
                    s = "!" + s
                elif re.match(r"\s*!\*+Mohammad MRT", s):
                    enable = 2
                elif enable:
                    s = s.replace("!", "", 1)
                    enable -= 1
        if n == 1481 and "WRITE(15,*)" in s:
            s = "!" + s
        s = re.sub(r"\bIF\(0\)THEN", "IF(.FALSE.)THEN", s)
        s = re.sub(r"\bIF\(1\)THEN", "IF(.TRUE.)THEN", s)
        if re.search(r";\s*PAUSE\s*$", s):
            s = re.sub(r";\s*PAUSE\s*$", "", s)
        s = re.sub(r"(INTEGER,PARAMETER :: NX=)\d+", r"\g<1>%d" % nx, s)
        s = re.sub(r"(INTEGER,PARAMETER :: NY=)\d+", r"\g<1>%d" % ny, s)
        s = re.sub(r"(LOGICAL,PARAMETER :: NON_NEWTONIAN=)\s*\.\w+\.",
                   r"\g<1> .%s." % ("TRUE" if non_newtonian else "FALSE"), s)
        s = re.sub(r"(LOGICAL,PARAMETER :: WRITE_OUTPUT=)\s*\.\w+\.",
                   r"\g<1> .%s." % ("TRUE" if write_output else "FALSE"), s)
        s = re.sub(r"(INTEGER,PARAMETER :: TIME_MAX=).*$",
                   r"\g<1>%d" % time_max, s)
        if uxn is not None and re.match(r"\s*UXN=0\.0;UYN=0D0", s):
            s = s.replace("UXN=0.0", "UXN=" + uxn, 1)
        if re.match(r"\s*SUBROUTINE Collision_MRT\(", s):
            in_collide = True
            in_rate = True
        elif re.match(r"\s*END SUBROUTINE Collision_MRT", s):
            in_rate = False
        if rheology == "casson" and in_rate:
            # the Casson lines F6 carries commented out next to the
            # Papanastasiou-Bingham ones (F6:939-945)
            if re.match(r"\s*Nu_A=.*!Bingham\s*$", s):
                s = "!" + s
            elif re.match(r"\s*!Nu_A=.*!Casson", s):
                s = s.replace("!", "", 1)
        out.append(s)
        if in_collide and re.match(r"\s*DO I=0,nX\s*$", s):
            if local_rate:
                out.append("\t\t\tCALL ComputeRelaxation_MRT(OMEGA_BN(I,J))")
            in_collide = False
    return "\n".join(out) + "\n"


def splice_boundary(text, src, uxn=None):
    """F6 with its BOUNDARYCONDITION (F6:1195-1390) replaced by the one of
    reference program src: F1's (F1:1222-1417: Zou-He edges, bounce-back
    corners) or F4's (F4:1205-1402: equilibrium edges and corners), with
    F6's argument list (neither routine uses KK).  A longer routine loses
    comment-only lines (the first ones after its declarations) so the
    line-numbered patches after it still apply.  uxn: top-wall velocity for
    F1's Zou-He lid (F1:1230, zero in F1)."""
    with open(src, "rb") as fh:
        other = fh.read().decode("ascii", errors="replace").replace("\r", "")
    sub = cut(other, r"\s*SUBROUTINE BOUNDARYCONDITION\(", r"\s*END SUBROUTINE BOUNDARYCONDITION\b")
    sub[0] = sub[0].replace("(F,F_POST,RHO,UX,UY)", "(F,F_POST,RHO,UX,UY,KK)")
    for k, line in enumerate(sub):
        if re.match(r"\s*INTEGER :: I,J\s*$", line):
            sub[k] = line.replace("I,J", "I,J,KK")
        if uxn is not None and re.match(r"\s*UXN=0\.0;UYN=0\.0\s*$", line):
            sub[k] = line.replace("UXN=0.0", "UXN=" + uxn, 1)
    lines = text.replace("\r", "").split("\n")
    a = next(k for k, l in enumerate(lines) if re.match(r"\s*SUBROUTINE BOUNDARYCONDITION\(", l))
    b = next(k for k in range(a, len(lines))
             if re.match(r"\s*END SUBROUTINE BOUNDARYCONDITION\b", lines[k]))
    extra = len(sub) - (b - a + 1)
    assert extra >= 0, "the spliced routine is shorter than F6's (line-numbered patches)"
    while extra:
        k = next(k for k in range(1, len(sub)) if re.match(r"\s*!", sub[k]))
        del sub[k]
        extra -= 1
    return "\n".join(lines[:a] + sub + lines[b + 1:]) + "\n"


def build(name, cfg, flags):
    with open(F6, "rb") as fh:
        text = fh.read().decode("ascii")
    cfg = dict(cfg)
    bnd = cfg.pop("boundary", None)
    if bnd == "zouhe":
        text = splice_boundary(text, F1, cfg.pop("zouhe_uxn", None))
    elif bnd == "equilibrium":
        text = splice_boundary(text, F4)
    src = patch_source(text, **cfg)
    os.makedirs(OUT_DIR, exist_ok=True)
    tmpdir = tempfile.mkdtemp(prefix="refbuild_", dir="/tmp")
    try:
        path = os.path.join(tmpdir, name + ".f90")
        with open(path, "w") as fh:
            fh.write(src)
        exe = os.path.join(OUT_DIR, name)
        cmd = [FLANG] + flags + [path, "-o", exe]
        subprocess.run(cmd, check=True, cwd=tmpdir,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return exe


def cut(text, start, end):
    """Lines from the first matching `start` through the next `end`."""
    lines = text.replace("\r", "").split("\n")
    out, on = [], False
    for line in lines:
        if not on and re.match(start, line):
            on = True
        if on:
            out.append(line)
            if re.match(end, line):
                return out
    raise RuntimeError("section %r not found" % start)


def build_f3_contact(flags):
    with open(F3, "rb") as fh:
        text = fh.read().decode("ascii", errors="replace")
    src = cut(text, r"\s*MODULE PARAM\b", r"\s*END MODULE PARAM\b")
    src += cut(text, r"\s*SUBROUTINE ComputeCollisionForce\(",
               r"\s*END SUBROUTINE ComputeCollisionForce\b")
    os.makedirs(OUT_DIR, exist_ok=True)
    tmpdir = tempfile.mkdtemp(prefix="refbuild_", dir="/tmp")
    try:
        path = os.path.join(tmpdir, "f3_contact_ref.f90")
        with open(path, "w") as fh:
            fh.write("\n".join(src) + "\n")
        exe = os.path.join(OUT_DIR, "f3_contact")
        cmd = [FLANG] + flags + [path, os.path.join(HERE, "f3_contact_driver.f90"), "-o", exe]
        subprocess.run(cmd, check=True, cwd=tmpdir,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return exe


# contact-model builds: name -> (source program, model switched on)
CONTACT_MODELS = {
    "f4_contact_wt": ("F4", "W-T"),
    "f4_contact_lj": ("F4", "L-J"),
    "f1_contact_wt": ("F1", "W-T"),
}


def set_contact_model(lines, model):
    """ComputeCollisionForce with the IF(0)/IF(1) switch that follows each
    model's banner comment ("L-J MODEL", "W-T MODEL", "GLOWINSKY MODEL")
    set to .TRUE. for `model` and .FALSE. for the others."""
    out, pending = [], None
    for line in lines:
        m = re.match(r"\s*!\*+(L-J|W-T|GLOWINSKY) MODEL", line)
        if m:
            pending = m.group(1)
        elif pending and re.match(r"\s*IF\((0|1)\)THEN", line):
            line = re.sub(r"IF\((0|1)\)THEN", "IF(.%s.)THEN" % ("TRUE" if pending == model else "FALSE"),
                          line, count=1)
            pending = None
        out.append(line)
    return out


def build_contact_model(name, flags):
    """MODULE PARAM and ComputeCollisionForce(DISK,WALL) of F4 or F1 with one
    contact model switched on, compiled with oracle/contact_model_driver.f90
    into oracle/_ref/<name>."""
    prog, model = CONTACT_MODELS[name]
    with open({"F4": F4, "F1": F1}[prog], "rb") as fh:
        text = fh.read().decode("ascii", errors="replace")
    src = cut(text, r"\s*MODULE PARAM\b", r"\s*END MODULE PARAM\b")
    routine = cut(text, r"\s*SUBROUTINE ComputeCollisionForce\(",
                  r"\s*END SUBROUTINE ComputeCollisionForce\b")
    src += set_contact_model(routine, model)
    os.makedirs(OUT_DIR, exist_ok=True)
    tmpdir = tempfile.mkdtemp(prefix="refbuild_", dir="/tmp")
    try:
        path = os.path.join(tmpdir, name + "_ref.f90")
        with open(path, "w") as fh:
            fh.write("\n".join(src) + "\n")
        exe = os.path.join(OUT_DIR, name)
        cmd = [FLANG] + flags + [path, os.path.join(HERE, "contact_model_driver.f90"), "-o", exe]
        subprocess.run(cmd, check=True, cwd=tmpdir,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return exe


# Multi-particle builds of whole reference programs: name -> source program,
# grid, steps and the disks.  F3 (N-particle sedimentation) reads its disks
# from 0_Particle_Initial_Position.txt, which the fixture generator
# (tests/golden/make_golden.py) writes into the run directory; F2 (the
# two-disk DKT program) sets its two centres in the code (F2:161-162), which
# the build replaces with the values given here.
NPARTICLE_BUILDS = {
    "f3_multi": dict(src="F3", nx=160, ny=160, n_particle=12, time_max=300),
    "f2_dkt": dict(src="F2", nx=120, ny=240, time_max=400,
                   centers=[("60.0D0", "150.0D0"), ("60.5D0", "129.75D0")]),
}


def patch_nparticle(text, nx, ny, time_max, n_particle=None, centers=None, clock=False,
                    write_output=True):
    """F3 or F2 made runnable and observable, without touching the
    algorithm:
      * IF(0)/IF(1) to logicals and PAUSE removed, as for F6;
      * PARAMETERs NX, NY (F3:8-9, F2:7-8), TIME_MAX (F3:37, F2:35) and
        WRITE_OUTPUT (F3:6, F2:5: the final state dump, F3:425-447; on
        unless `write_output` is false); F3's
        initial N_PARTICLE (F3:68);
      * F3: the STOP after the initial domain image (F3:202), which ends the
        program before its time loop, commented out;
      * F2: the two initial centres (F2:161-162) set to `centers`;
      * two output lines: the run constants OMEGA, NU, A_G, RHO_S, RHO_F, PI
        written to ref_constants.txt after OMEGA is set (F3:99, F2:97), and,
        at the end of every loop iteration (after the mean-density line,
        F3:378, F2:371), KK and each disk's CENTER, VELOCITY, FORCE and
        ANGLE written to ref_history.txt; with `clock`, also KK and the
        program's CPU clock since START (F3:136, F2:138) to fort.98, so a
        timing can leave the set-up out."""
    lines = text.replace("\r", "").split("\n")
    out = []
    in_loop = False
    after_image = False
    stop_done = False
    for line in lines:
        s = line
        s = re.sub(r"\bIF\(0\)THEN", "IF(.FALSE.)THEN", s)
        s = re.sub(r"\bIF\(1\)THEN", "IF(.TRUE.)THEN", s)
        s = re.sub(r";\s*PAUSE\s*$", "", s)
        s = re.sub(r"(INTEGER,PARAMETER :: NX=)\d+", r"\g<1>%d" % nx, s)
        s = re.sub(r"(INTEGER,PARAMETER :: NY=)\d+", r"\g<1>%d" % ny, s)
        s = re.sub(r"(INTEGER,PARAMETER :: TIME_MAX=).*$", r"\g<1>%d" % time_max, s)
        s = re.sub(r"(LOGICAL,PARAMETER :: WRITE_OUTPUT=)\s*\.\w+\.",
                   r"\g<1> .%s." % ("TRUE" if write_output else "FALSE"), s)
        if n_particle is not None:
            s = re.sub(r"(INTEGER :: N_PARTICLE=)\d+", r"\g<1>%d" % n_particle, s)
        if centers is not None:
            m = re.match(r"(\s*)DISK\(([12])\)%CENTER=\(/", s)
            if m:
                x, y = centers[int(m.group(2)) - 1]
                s = "%sDISK(%s)%%CENTER=(/%s,%s/)" % (m.group(1), m.group(2), x, y)
        if n_particle is not None and after_image and not stop_done and re.match(r"\s*STOP\s*$", s):
            s = "!" + s
            stop_done = True
        if re.match(r"\s*CALL WRITE_DOMAINIMAGE\(DOMAIN_IMAGE,0,10\)", s):
            after_image = True
        out.append(s)
        if re.match(r"\s*OMEGA=1\.0/\(3\.\*NU\+0\.5\)", s):
            out.append("\tOPEN(96,FILE='ref_constants.txt');"
                       "WRITE(96,*) OMEGA,NU,A_G,RHO_S,RHO_F,PI;CLOSE(96)")
            out.append("\tOPEN(97,FILE='ref_history.txt')")
        if re.match(r"\s*DO WHILE\(KK<TIME_MAX\)", s):
            in_loop = True
        if in_loop and re.match(r"\s*WRITE\(8,\*\) KK,AvgDen", s):
            out.append("\t\tWRITE(97,*) KK,(DISK(IP)%CENTER,DISK(IP)%VELOCITY,"
                       "DISK(IP)%FORCE,DISK(IP)%ANGLE,IP=1,N_PARTICLE)")
            if clock:
                out.append("\t\tCALL CPU_TIME(FINISH);WRITE(98,*) KK,FINISH-START;FLUSH(98)")
            in_loop = False
    assert n_particle is None or stop_done, "F3's STOP before the time loop not found"
    return "\n".join(out) + "\n"


def build_nparticle(name, cfg, flags):
    cfg = dict(cfg)
    src_path = {"F3": F3, "F2": F2}[cfg.pop("src")]
    with open(src_path, "rb") as fh:
        text = fh.read().decode("ascii", errors="replace")
    src = patch_nparticle(text, **cfg)
    os.makedirs(OUT_DIR, exist_ok=True)
    tmpdir = tempfile.mkdtemp(prefix="refbuild_", dir="/tmp")
    try:
        path = os.path.join(tmpdir, name + ".f90")
        with open(path, "w") as fh:
            fh.write(src)
        exe = os.path.join(OUT_DIR, name)
        r = subprocess.run([FLANG] + flags + [path, "-o", exe], cwd=tmpdir,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        if r.returncode:
            raise RuntimeError("%s build failed:\n%s" % (name, r.stdout[-3000:]))
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return exe


SHIM_ROUTINES = ["Collision_MRT", "STREAMING", "BOUNDARYCONDITION", "MACROSCOPICVARIABLES",
                 "ComputeConcenterationFunction", "ComputeForceField_SPM",
                 "ComputeForce_HYDRO", "VelocityCorrector", "ComputeCollisionForce"]
# variant -> extra preprocessor definitions of the shim
SHIM_VARIANTS = {
    "f6_bing_lagged": [],
    "f6_newt_lid": ["-DLBM_SHIM_UXN=0.01D0"],
    "f6_newt_mohammad": [],
}
PKG = os.path.join(os.path.dirname(HERE), "2d_nnf_lbm_amd")


def drop_routines(text, names):
    """The source without the listed subroutines (SUBROUTINE name( through
    its END SUBROUTINE or bare END line)."""
    out, skip = [], False
    for line in text.split("\n"):
        if not skip:
            m = re.match(r"\s*SUBROUTINE\s+(\w+)\s*\(", line, re.I)
            if m and m.group(1).lower() in [n.lower() for n in names]:
                skip = True
                continue
            out.append(line)
        elif re.match(r"\s*END(\s+SUBROUTINE\b.*)?\s*(!.*)?$", line, re.I):
            skip = False
    return "\n".join(out) + "\n"


def build_shim(name, cfg, defines, flags):
    lib = os.path.join(PKG, "liblbm_amd.so")
    if not os.path.exists(lib):
        raise RuntimeError("build liblbm_amd.so before the shim programs")
    with open(F6, "rb") as fh:
        text = fh.read().decode("ascii")
    src = drop_routines(patch_source(text, **cfg), SHIM_ROUTINES)
    tmpdir = tempfile.mkdtemp(prefix="refbuild_", dir="/tmp")
    try:
        path = os.path.join(tmpdir, name + "_main.f90")
        with open(path, "w") as fh:
            fh.write(src)
        exe = os.path.join(OUT_DIR, "f6_shim_" + name[3:])
        cmd = ([FLANG] + flags + ["-cpp"] + defines + ["-J", tmpdir,
               os.path.join(PKG, "fortran", "lbm_amd.f90"), path,
               os.path.join(PKG, "fortran", "lbm_f6_shim.F90"),
               "-L" + PKG, "-llbm_amd", "-Wl,-rpath,$ORIGIN/../../2d_nnf_lbm_amd", "-o", exe])
        r = subprocess.run(cmd, cwd=tmpdir, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                           text=True)
        if r.returncode:
            raise RuntimeError("shim build failed:\n" + r.stdout[-3000:])
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return exe


def main(argv):
    if not os.path.exists(F6):
        print("build_ref: reference source not present; nothing to build")
        return 0
    parity_flags = ["-O2", "-ffp-contract=off"]
    timing_flags = ["-O3", "-march=x86-64-v3"]
    built = []
    for name, cfg in VARIANTS.items():
        built.append(build(name, cfg, parity_flags))
    if os.path.exists(F3):
        built.append(build_f3_contact(parity_flags + ["-fno-automatic"]))
    for name, cfg in NPARTICLE_BUILDS.items():
        if os.path.exists({"F3": F3, "F2": F2}[cfg["src"]]):
            built.append(build_nparticle(name, cfg, parity_flags + ["-fno-automatic"]))
    if os.path.exists(F4) and os.path.exists(F1):
        for name in CONTACT_MODELS:
            built.append(build_contact_model(name, parity_flags))
    if os.path.exists(os.path.join(PKG, "liblbm_amd.so")):
        for name, defines in SHIM_VARIANTS.items():
            built.append(build_shim(name, VARIANTS[name], defines, parity_flags))
    if "--timing" in argv or "--all" in argv:
        for name, cfg in TIMING_VARIANTS.items():
            built.append(build(name, cfg, timing_flags))
        if os.path.exists(F4):
            for name, (side, steps) in F4_TIMING.items():
                built.append(build_f4_timing(name, side, steps, timing_flags))
        if os.path.exists(F3):
            # its static arrays exceed 2 GB: the medium code model
            for name, cfg in F3_TIMING.items():
                built.append(build_nparticle(name, cfg, timing_flags + ["-mcmodel=medium"]))
    for b in built:
        print("built", os.path.relpath(b))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
