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
# usage: oracle/ref_build.sh [variant ...]     (default: all variants)
set -euo pipefail
REF=${J1J2_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
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
hot_init_l6h4 L=6 H=4 SAMPLERMODE='PT' NREPLICA=2 TMIN=8.0D0 TMAX=11.5D0 START='HOT' NEQ=0 NMESS=0 NCONF=2 IRINIT=7 WRITEBINDTIME=.false.
str_init_l12  L=12 H=10 SAMPLERMODE='PT' NREPLICA=2 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='Y' NEQ=0 NMESS=0 NCONF=1 IRINIT=11 WRITEBINDTIME=.false.
pt_l8_stat    L=8 H=8 SAMPLERMODE='PT' NREPLICA=4 TMIN=8.0D0 TMAX=11.5D0 START='STR' STRAXIS='X' NEQ=2000 NMESS=40000 NCONF=1 IRINIT=1 WRITEBINDTIME=.false. WRITELOCNEM=.false.
local_l8      L=8 H=8 SAMPLERMODE='LOCAL' TMIN=10.0D0 TMAX=10.0D0 START='HOT' NEQ=1000 NMESS=1000 NCONF=1 IRINIT=1 WRITEBINDTIME=.false.
local_frozen  L=8 H=8 SAMPLERMODE='LOCAL' TMIN=0.01D0 TMAX=0.01D0 START='FER' phi=0.0D0 NEQ=10 NMESS=200 NCONF=1 IRINIT=1 WRITEBINDTIME=.false.
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
    ( cd "$tmp" && "$FC" -cpp -O2 -w -I"$MPI_INC" \
        "$REF/sim/corr_rand_3d.f90" "$src" \
        "$REF/vendor/fft2d/fftsg3d.f" "$REF/vendor/fft2d/fftsg.f" \
        -L"$MPI_LIB" -Wl,-rpath,"$MPI_LIB" -lmpifort -lmpi -o "$OUT/j1j2_$name" )
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
