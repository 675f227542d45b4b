#!/bin/bash
# Builds the reference sam(oa)^2 SWE binary (TEST INFRASTRUCTURE ONLY).
#
# Equivalent of SCons `scenario=swe flux_solver=aug_riemann precision=double
# asagi=No mpi=nompi openmp=...` (SConstruct:143-356) without running SCons:
# every source of src/SConscript:14-100 is compiled with amdflang, module
# order resolved by repeated passes.  The sources are copied to a scratch
# directory under /tmp (never into this repository) where two documented
# compiler workarounds are applied:
#   1. Conformity/Conformity.f90:47-51 is compiled with -D__GFORTRAN__ (flang
#      rejects the logical-indexed parameter array of the other branch).
#   2. SFC_traversal.f90:151,165 call the GNU `signal` extension that flang
#      lacks; the two calls are replaced by `i_status = 0` (only the SIGTERM
#      soft exit is lost).
# VARIANT=bath additionally applies two one-line scenario patches, so that the
# oracle also exercises bathymetry and reflecting walls:
#   SWE_initialize.f90:185   bathymetry = 0  ->  b = -0.5 + 0.25 x + 0.1 cos(2 y)
#                            (a beach rising towards +x, dry for x > ~2)
#   SWE_euler_timestep.f90:248 OUTFLOW copy -> mirror ghost p - 2 (p.n) n
# VARIANT=dam raises the dam break of SWE_initialize.f90:354 from (1 m | dry)
# to (10 m | 5 m), a wet shock strong enough to trigger refinement while the
# mesh adapts every step (-nadapt 1):
#   SWE_initialize.f90:354   hL = 1, hR = 0  ->  hL = 10, hR = 5
# VARIANT=radial is the radial dam break in a closed basin: the unit square
# (no offset), eta = 10 m inside the circle of radius 0.1 around (0.5, 0.5)
# and 5 m outside, a flat bed and reflecting walls:
#   SWE.f90:169-170          scaling = 10, offset = -5 (5, 5)  ->  scaling = 1, offset = 0
#   SWE_initialize.f90:354   hL = 1, hR = 0  ->  hL = 10, hR = 5
#   SWE_initialize.f90:362   xs(1) < 0  ->  norm2(xs - 0.5) < 0.1
#   SWE_euler_timestep.f90:248 OUTFLOW copy -> mirror ghost p - 2 (p.n) n
# FLUX=... builds one of the reference's other flux solvers (see below).
# Only the linked binary is written to $OUT (oracle/_ref/).
set -euo pipefail
: "${REF:?}" "${REFTMP:?}" "${OUT:?}"
VARIANT="${VARIANT:-plain}"
rm -rf "$REFTMP"
mkdir -p "$REFTMP/mod" "$REFTMP/obj"
cp -r "$REF/src" "$REFTMP/src"
sed -i 's/^\(\s*\)i_status = signal(SIGTERM, signal_terminate)$/\1i_status = 0/; s/^\(\s*\)i_status = signal(SIGTERM, 0)$/\1i_status = 0/' "$REFTMP/src/SFC_traversal.f90"
if [ "$VARIANT" = "bath" ]; then
    sed -i '185s/^\(\s*\)bathymetry = 0.0_SR$/\1bathymetry = -0.5_SR + 0.25_SR * xs(1) + 0.1_SR * cos(2.0_SR * xs(2))/' "$REFTMP/src/SWE/SWE_initialize.f90"
    sed -i '248s/^\(\s*\)bnd_rep = rep%Q(1)$/\1bnd_rep = t_state(rep%Q(1)%h, rep%Q(1)%p - 2.0_GRID_SR * dot_product(rep%Q(1)%p, edge%transform_data%normal) * edge%transform_data%normal, rep%Q(1)%b)/' "$REFTMP/src/SWE/SWE_euler_timestep.f90"
    grep -q "0.1_SR \* cos(2.0_SR" "$REFTMP/src/SWE/SWE_initialize.f90"
    grep -q "2.0_GRID_SR \* dot_product" "$REFTMP/src/SWE/SWE_euler_timestep.f90"
elif [ "$VARIANT" = "dam" ]; then
    sed -i '354s/hL = 1.0_SR, hR = 0.0_SR/hL = 10.0_SR, hR = 5.0_SR/' "$REFTMP/src/SWE/SWE_initialize.f90"
    grep -q "hL = 10.0_SR, hR = 5.0_SR" "$REFTMP/src/SWE/SWE_initialize.f90"
elif [ "$VARIANT" = "radial" ]; then
    sed -i '169s/cfg%scaling = 10.0_GRID_SR$/cfg%scaling = 1.0_GRID_SR/' "$REFTMP/src/SWE/SWE.f90"
    sed -i '170s/cfg%offset = cfg%scaling \* \[-0.5_GRID_SR, -0.5_GRID_SR\]$/cfg%offset = [0.0_GRID_SR, 0.0_GRID_SR]/' "$REFTMP/src/SWE/SWE.f90"
    sed -i '354s/hL = 1.0_SR, hR = 0.0_SR/hL = 10.0_SR, hR = 5.0_SR/' "$REFTMP/src/SWE/SWE_initialize.f90"
    sed -i '362s/if (xs(1) < 0.0_SR) then/if (norm2(xs - 0.5_SR) < 0.1_SR) then/' "$REFTMP/src/SWE/SWE_initialize.f90"
    sed -i '248s/^\(\s*\)bnd_rep = rep%Q(1)$/\1bnd_rep = t_state(rep%Q(1)%h, rep%Q(1)%p - 2.0_GRID_SR * dot_product(rep%Q(1)%p, edge%transform_data%normal) * edge%transform_data%normal, rep%Q(1)%b)/' "$REFTMP/src/SWE/SWE_euler_timestep.f90"
    grep -q "cfg%scaling = 1.0_GRID_SR" "$REFTMP/src/SWE/SWE.f90"
    grep -q "cfg%offset = \[0.0_GRID_SR, 0.0_GRID_SR\]" "$REFTMP/src/SWE/SWE.f90"
    grep -q "hL = 10.0_SR, hR = 5.0_SR" "$REFTMP/src/SWE/SWE_initialize.f90"
    grep -q "norm2(xs - 0.5_SR) < 0.1_SR" "$REFTMP/src/SWE/SWE_initialize.f90"
    grep -q "2.0_GRID_SR \* dot_product" "$REFTMP/src/SWE/SWE_euler_timestep.f90"
elif [ "$VARIANT" != "plain" ]; then
    echo "build_samoa.sh: unknown VARIANT $VARIANT" >&2
    exit 1
fi
# FLUX selects the flux solver as SCons flux_solver= does (SConstruct:224-237):
# AUG_RIEMANN (default), FWAVE, LF, LLF, LF_BATH or LLF_BATH.
FLUX="${FLUX:-AUG_RIEMANN}"
case "$FLUX" in AUG_RIEMANN|FWAVE|LF|LLF|LF_BATH|LLF_BATH) ;; *) echo "build_samoa.sh: unknown FLUX $FLUX" >&2; exit 1;; esac
# OPT: optimisation flags.  The parity builds keep every rounding as written
# (-O2 -ffp-contract=off); the timing builds take SConstruct:306-310's
# release flags (-O3, or -Ofast -march=native).
OPT="${OPT:--O2 -ffp-contract=off}"
FL="-cpp -D_SWE -D_${FLUX}_FLUX -D_DOUBLE_PRECISION -D_ADAPT_INTEGRATE -D_DARCY_LAYERS=0 -D_DEBUG_LEVEL=1 -D_DARCY_MOB_QUADRATIC -D_PERM_MEAN_GEOMETRIC -fopenmp $OPT -I$REFTMP/src -I$REFTMP/src/Samoa -module-dir $REFTMP/mod -I$REFTMP/mod"
SRCS=$(python3 -c "
import re
t = open('$REF/src/SConscript').read()
print(' '.join(re.findall(r'\"([^\"]+\.f90)\"', t) + re.findall(r'\"([^\"]+\.f)\"', t)))
")
left="$SRCS"
for pass in $(seq 1 40); do
    next=""
    for f in $left; do
        o="$REFTMP/obj/$(echo "$f" | tr '/' '_').o"
        xf=""
        [ "$f" = "Conformity/Conformity.f90" ] && xf="-D__GFORTRAN__"
        if ! amdflang $FL $xf -c "$REFTMP/src/$f" -o "$o" > "$o.log" 2>&1; then next="$next $f"; fi
    done
    [ -z "$next" ] && break
    if [ "$next" = "$left" ]; then
        echo "build_samoa.sh: cannot compile:$next" >&2
        for f in $next; do grep -m3 error "$REFTMP/obj/$(echo "$f" | tr '/' '_').o.log" >&2 || true; done
        exit 1
    fi
    left="$next"
done
amdflang -fopenmp "$REFTMP"/obj/*.o -o "$OUT.tmp"
mv "$OUT.tmp" "$OUT"
echo "build_samoa.sh: built $OUT"
