#!/usr/bin/env bash
# Builds the unmodified reference program (sim/J1J2_6tv.f90 + sim/corr_rand_3d.f90
# + vendor/fft2d/fftsg3d.f + vendor/fft2d/fftsg.f) from its sources under
# /root/reference, the way its README builds it (mpif90 -cpp, here amdflang +
# the image's MPICH), into oracle/_ref/.  Every run parameter of the reference is
# a compile-time constant (J1J2_6tv.f90:43-96), so each variant below sets the
# parameter block -- and nothing else -- in a scratch copy under a mktemp
# directory outside the repository, which is deleted after compiling.  Only the
# binaries land in oracle/_ref/ (git-ignored).  Test infrastructure only.
#
# Variants marked DROPIN=1 are the reference program with the engine dropped
# in, as a maintainer would: in the scratch copy the PT SERIAL calls of
# Initialize_pt_state, Pt_equilibration_loop and Pt_measurement_loop
# (J1J2_6tv.f90:340-342) become one call of Gpu_pt_run, which
# j1j2-public_amd/fortran/gpu_pt_run.inc provides after the program's
# `contains` (J1J2_6tv.f90:411); linked with the iso_c_binding module and
# libj1j2_engine.so (build it first: make -C j1j2-public_amd).
#
# usage: oracle/ref_build.sh [variant ...]     (default: all variants)
set -euo pipefail
REF=${J1J2_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
PKG=$(cd "$HERE/../j1j2-public_amd" && pwd)
FC=${FC:-/opt/rocm/llvm/bin/amdflang}
MPI_INC=${MPI_INC:-/opt/conda/include}
MPI_LIB=${MPI_LIB:-/opt/conda/lib}

if [ ! -f "$REF/sim/J1J2_6tv.f90" ]; then
    echo "ref_build: reference sources not found under $REF; nothing built" >&2
    exit 0
fi
mkdir -p "$OUT"

# name | KEY=VALUE ... (applied to the parameter block only)
variants() {
cat <<'EOF'
hot_init_l8   L=8 H=8 SAMPLERMODE='PT' NREPLICA=4 TMIN=8.0D0 TMAX=11.5D0 START='HOT' NEQ=0 NMESS=0 NCONF=3 IRINIT=1 WRITEBINDTIME=.false.
hot_init_l6h4 L=6 H=4 SAMPLERMODE='PT' NREPLICA=2 TMIN=8.0D0 TMAX=11.5D0 START='HOT' NEQ=0 NMESS=0 NCONF=2 IRINIT=7 WRITEBINDTIME=.false. WRITENUMFLIPS=.true.
str_init_l12  L=12 H=10 SAMPLERMODE='PT' NREPLICA=2 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='Y' NEQ=0 NMESS=0 NCONF=1 IRINIT=11 WRITEBINDTIME=.false.
pt_l8_stat    L=8 H=8 SAMPLERMODE='PT' NREPLICA=4 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='X' NEQ=2000 NMESS=40000 NCONF=1 IRINIT=1 WRITEBINDTIME=.false. WRITELOCNEM=.false.
local_l8      L=8 H=8 SAMPLERMODE='LOCAL' TMIN=10.0D0 TMAX=10.0D0 START='HOT' NEQ=1000 NMESS=1000 NCONF=1 IRINIT=1 WRITEBINDTIME=.false.
local_frozen  L=8 H=8 SAMPLERMODE='LOCAL' TMIN=0.01D0 TMAX=0.01D0 START='FER' phi=0.0D0 NEQ=10 NMESS=200 NCONF=1 IRINIT=1 WRITEBINDTIME=.false.
pt_frozen     L=8 H=8 SAMPLERMODE='PT' NREPLICA=5 TMIN=0.01D0 TMAX=0.02D0 START='FER' phi=0.0D0 NEQ=11 NMESS=7 NCONF=1 IRINIT=1 WRITEBINDTIME=.false.
pt_l16_c2     L=16 H=16 SAMPLERMODE='PT' NREPLICA=16 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='X' NEQ=2000 NMESS=2000 NCONF=1 IRINIT=1 WRITEBINDTIME=.false. WRITELOCNEM=.false. WRITEPHITABLE=.false.
pt_dyn_l6h4   L=6 H=4 SAMPLERMODE='PT' NREPLICA=3 TMIN=8.0D0 TMAX=11.5D0 START='HOT' NEQ=100 NMESS=300 NCONF=2 IRINIT=7 WRITEBINDTIME=.false. WRITENUMFLIPS=.true.
local_dyn_l6h4 L=6 H=4 SAMPLERMODE='LOCAL' TEMPPROTOCOL='INDEP' TMIN=4.0D0 TMAX=8.0D0 DT=2.0D0 START='HOT' NEQ=100 NMESS=200 NCONF=2 IRINIT=5 WRITEBINDTIME=.false. WRITENUMFLIPS=.true.
local_indep_init L=6 H=4 SAMPLERMODE='LOCAL' TEMPPROTOCOL='INDEP' TMIN=4.0D0 TMAX=8.0D0 DT=2.0D0 START='HOT' NEQ=0 NMESS=0 NCONF=2 IRINIT=5 WRITEBINDTIME=.false.
local_carry_init L=6 H=4 SAMPLERMODE='LOCAL' TEMPPROTOCOL='CARRY' TMIN=4.0D0 TMAX=8.0D0 DT=2.0D0 START='HOT' NEQ=0 NMESS=0 NCONF=2 IRINIT=5 WRITEBINDTIME=.false.
dropin_hot_init_l8 DROPIN=1 L=8 H=8 SAMPLERMODE='PT' NREPLICA=4 TMIN=8.0D0 TMAX=11.5D0 START='HOT' NEQ=0 NMESS=0 NCONF=3 IRINIT=1 WRITEBINDTIME=.false.
dropin_pt_frozen   DROPIN=1 L=8 H=8 SAMPLERMODE='PT' NREPLICA=5 TMIN=0.01D0 TMAX=0.02D0 START='FER' phi=0.0D0 NEQ=11 NMESS=7 NCONF=1 IRINIT=1 WRITEBINDTIME=.false.
dropin_pt_dyn      DROPIN=1 L=8 H=8 SAMPLERMODE='PT' NREPLICA=4 TMIN=8.0D0 TMAX=11.5D0 START='HOT' NEQ=30 NMESS=60 NCONF=3 IRINIT=1 WRITEBINDTIME=.false. WRITENUMFLIPS=.true.
cpu_l32       L=32 H=32 SAMPLERMODE='PT' NREPLICA=16 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='X' NEQ=100 NMESS=0 NCONF=16 IRINIT=1 WRITEBINDTIME=.false. WRITELOCNEM=.false. WRITEHIST=.false.
cpu_l16       L=16 H=16 SAMPLERMODE='PT' NREPLICA=16 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='X' NEQ=1000 NMESS=0 NCONF=1 IRINIT=1 WRITEBINDTIME=.false. WRITELOCNEM=.false. WRITEHIST=.false.
EOF
}

build_one() {
    local name=$1; shift
    local tmp
    tmp=$(mktemp -d /tmp/j1j2ref.XXXXXX)
    local src="$tmp/J1J2_6tv.f90"
    local script=""
    local dropin=0
    local params=()
    for kv in "$@"; do
        if [ "$kv" = "DROPIN=1" ]; then dropin=1; else params+=("$kv"); fi
    done
    set -- "${params[@]}"
    for kv in "$@"; do
        local key=${kv%%=*} val=${kv#*=}
        # lines 40-96 hold the parameter block; replace the value after "KEY[*n] ="
        script+="40,96s/^\\([^!]*parameter *:: *${key}\\(\\*[0-9]*\\)\\{0,1\\} *= *\\)[^!]*/\\1${val} /;"
    done
    sed -e "$script" "$REF/sim/J1J2_6tv.f90" > "$src"
    local block
    block=$(sed -n '40,96p' "$src")
    for kv in "$@"; do
        local key=${kv%%=*}
        if ! grep -q "parameter *:: *${key}\(\*[0-9]*\)\{0,1\} *= *${kv#*=} " <<< "$block"; then
            echo "ref_build: failed to set $key for $name" >&2; rm -rf "$tmp"; exit 1
        fi
    done
    if [ $dropin = 0 ]; then
        ( cd "$tmp" && "$FC" -cpp -O2 -w -I"$MPI_INC" \
            "$REF/sim/corr_rand_3d.f90" "$src" \
            "$REF/vendor/fft2d/fftsg3d.f" "$REF/vendor/fft2d/fftsg.f" \
            -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpifort -lmpi -o "$OUT/j1j2_$name" )
    else
        if [ ! -f "$PKG/libj1j2_engine.so" ]; then
            echo "ref_build: $name needs $PKG/libj1j2_engine.so (make -C j1j2-public_amd)" >&2; rm -rf "$tmp"; exit 1
        fi
        if ! { sed -n '340p' "$src" | grep -q '^ *call Initialize_pt_state$' &&
               sed -n '341p' "$src" | grep -q '^ *call Pt_equilibration_loop$' &&
               sed -n '342p' "$src" | grep -q '^ *call Pt_measurement_loop$' &&
               sed -n '411p' "$src" | grep -q '^ *contains$'; }; then
            echo "ref_build: $name: the PT SERIAL calls or contains are not where expected" >&2; rm -rf "$tmp"; exit 1
        fi
        sed -i -e '340s/call Initialize_pt_state/call Gpu_pt_run/' -e '341,342d' \
            -e "411a\\      include 'gpu_pt_run.inc'" "$src"
        ( cd "$tmp" && "$FC" -cpp -O2 -w -I"$MPI_INC" -I"$PKG/fortran" -module-dir "$tmp" \
            "$PKG/fortran/j1j2_engine_mod.f90" "$REF/sim/corr_rand_3d.f90" "$src" \
            "$REF/vendor/fft2d/fftsg3d.f" "$REF/vendor/fft2d/fftsg.f" \
            -L"$PKG" -lj1j2_engine -Wl,-rpath,'$ORIGIN/../../j1j2-public_amd' \
            -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpifort -lmpi -o "$OUT/j1j2_$name" )
    fi
    rm -rf "$tmp"
    echo "ref_build: built $OUT/j1j2_$name"
}

want=("$@")
while read -r name rest; do
    [ -z "$name" ] && continue
    if [ ${#want[@]} -gt 0 ]; then
        keep=0; for w in "${want[@]}"; do [ "$w" = "$name" ] && keep=1; done
        [ $keep = 1 ] || continue
    fi
    # shellcheck disable=SC2086
    build_one "$name" $rest
done < <(variants)
