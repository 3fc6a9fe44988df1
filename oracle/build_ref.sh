#!/bin/bash
# Builds the reference Fortran time-stepping path (oracle/_ref), a checker
# for the CPU oracle.  The reference modules on the path -- parameters,
# constants, error, derivs, variables, ic, solve -- are compiled UNMODIFIED
# from /root/reference (through the examples/ring symlinks, whose directory
# holds no parameters.in, so the compile-time configuration written here is
# found on the include path, as the reference's own setup script intends).
# mpif.h and the MPI library are the image's MPICH (/opt/conda).  The
# drivers are oracle/ref/ref_driver.f90, ref_diag.f90 and ref_ic.f90 (this
# project's harnesses); ref_ic is linked once per initial condition of
# oracle/ref/ic_cases.txt, each with that expression as the ic.in the
# reference includes.  Unused code (the FFT routine and the FFT diagnostics)
# is dropped at link time by LTO, so the absent FFTW2 library is never
# referenced: random_phase, which needs it, is not built.
# Outputs go to oracle/_ref/<nx>x<ny>x<nz>/{ref_driver,ref_diag} and
# oracle/_ref/ic/<case>/ref_ic only.
set -euo pipefail
REF=${GPE_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT=$HERE/_ref
FC=${FC:-/opt/rocm/lib/llvm/bin/amdflang}
MPI=${MPI_HOME:-/opt/conda}
[ -d "$REF/examples/ring" ] || { echo "build_ref.sh: $REF not present, skipping"; exit 0; }

write_pin() {
  local dir=$1 nx=$2 ny=$3 nz=$4 npy=${5:-1} npz=${6:-1}
  cat > "$dir/parameters.in" <<PIN
  integer, parameter :: pr = selected_real_kind(15,307)
  integer, parameter :: nyprocs = $npy
  integer, parameter :: nzprocs = $npz
  integer, parameter :: nx      = $nx
  integer, parameter :: ny      = $ny
  integer, parameter :: nz      = $nz
  type :: line_param
    real (pr)    :: x0, y0, z0, amp1, amp2, ll, sgn
    character(1) :: dir
    logical      :: imprint_phase
  end type line_param
  type (line_param), parameter :: vl1 = line_param(0.0_pr, 0.0_pr, 0.0_pr, &
    0.0_pr, 0.0_pr, 0.0_pr, 1.0_pr, 'z', .false.)
  type :: ring_param
    character(2) :: plane
    real (pr)    :: x0, y0, z0, r0, amp
    integer      :: mm
    real (pr)    :: r1
    integer      :: kk
    real (pr)    :: dir
  end type ring_param
  type (ring_param), parameter :: vr1 = ring_param('xy', 0.0_pr, 0.0_pr, &
    0.0_pr, 8.0_pr, 0.0_pr, 5, 0.0_pr, 10, 1.0_pr)
PIN
}

build_grid() {
  local nx=$1 ny=$2 nz=$3
  local dir=$OUT/${nx}x${ny}x${nz}
  mkdir -p "$dir"
  write_pin "$dir" $nx $ny $nz
  # the run starts from a restart file, so the compile-time IC is a constant
  echo "      init_cond = cmplx(1.0_pr, 0.0_pr, pr)" > "$dir/ic.in"
  local flags="-O2 -flto -I$dir -I$MPI/include -module-dir $dir"
  for m in parameters constants error derivs ic variables solve; do
    $FC $flags -c "$REF/examples/ring/$m.f90" -o "$dir/$m.o"
  done
  $FC $flags -c "$HERE/ref/ref_driver.f90" -o "$dir/ref_driver.o"
  $FC -O2 -flto -o "$dir/ref_driver" "$dir"/{parameters,constants,error,derivs,ic,variables,solve,ref_driver}.o \
      -L$MPI/lib -Wl,-rpath,$MPI/lib -lmpifort -lmpi -flto -fuse-ld=lld
  echo "built $dir/ref_driver"
  # diagnostics harness (momentum, velocity PDF, velocity correlation)
  $FC $flags -c "$HERE/ref/ref_diag.f90" -o "$dir/ref_diag.o"
  $FC -O2 -flto -o "$dir/ref_diag" "$dir"/{parameters,constants,error,derivs,ic,variables,ref_diag}.o \
      -L$MPI/lib -Wl,-rpath,$MPI/lib -lmpifort -lmpi -flto -fuse-ld=lld
  echo "built $dir/ref_diag"

  # Drop-in check: the same harness and the reference's own parameters,
  # constants, error, ic and variables modules, with solve and derivs
  # replaced by the GPU modules of gpe_amd/fortran (linked to libgpehip.so).
  local gdir=$dir/dropin
  local F=$HERE/../gpe_amd/fortran
  local LIBDIR=$(cd "$HERE/../gpe_amd" && pwd)
  mkdir -p "$gdir"
  local gflags="-O2 -flto -I$dir -I$MPI/include -module-dir $gdir -I$gdir"
  for m in parameters constants error; do
    $FC $gflags -c "$REF/examples/ring/$m.f90" -o "$gdir/$m.o"
  done
  $FC $gflags -c "$F/gpe_hip_c.f90" -o "$gdir/gpe_hip_c.o"
  $FC $gflags -c "$F/gpe_context.f90" -o "$gdir/gpe_context.o"
  $FC $gflags -c "$F/derivs.f90" -o "$gdir/derivs.o"
  $FC $gflags -c "$REF/examples/ring/ic.f90" -o "$gdir/ic.o"
  $FC $gflags -c "$REF/examples/ring/variables.f90" -o "$gdir/variables.o"
  $FC $gflags -c "$F/solve.f90" -o "$gdir/solve.o"
  $FC $gflags -c "$HERE/ref/ref_driver.f90" -o "$gdir/ref_driver.o"
  $FC -O2 -flto -o "$gdir/ref_driver_gpu" "$gdir"/{parameters,constants,error,gpe_hip_c,gpe_context,derivs,ic,variables,solve,ref_driver}.o \
      -L$LIBDIR -lgpehip -Wl,-rpath,$LIBDIR -L$MPI/lib -Wl,-rpath,$MPI/lib -lmpifort -lmpi -fuse-ld=lld
  echo "built $gdir/ref_driver_gpu"
}

# Initial conditions: the reference's ic module compiled once per case of
# ref/ic_cases.txt, with that case's expression as its ic.in, linked with
# the harness ref/ref_ic.f90.  Outputs: _ref/ic/<case>/ref_ic.
build_ics() {
  local nx=16 ny=12 nz=10
  local dir=$OUT/ic
  mkdir -p "$dir"
  write_pin "$dir" $nx $ny $nz
  # placeholder for the module build; each case compiles ic.f90 again
  echo "      init_cond = cmplx(1.0_pr, 0.0_pr, pr)" > "$dir/ic.in"
  local flags="-O2 -flto -I$dir -I$MPI/include -module-dir $dir"
  for m in parameters constants error derivs ic variables; do
    $FC $flags -c "$REF/examples/ring/$m.f90" -o "$dir/$m.o"
  done
  $FC $flags -c "$HERE/ref/ref_ic.f90" -o "$dir/ref_ic.o"
  local name expr
  while IFS='|' read -r name expr; do
    case "$name" in ''|'#'*) continue ;; esac
    local cdir=$dir/$name
    mkdir -p "$cdir"
    echo "  init_cond = $expr" > "$cdir/ic.in"
    # the case's ic.in comes first on the include path; the module file
    # is the same for every case and stays in $dir
    $FC -O2 -flto -I"$cdir" -I$dir -I$MPI/include -module-dir "$cdir" -c "$REF/examples/ring/ic.f90" -o "$cdir/ic.o"
    $FC -O2 -flto -o "$cdir/ref_ic" "$dir"/{parameters,constants,error,derivs}.o "$cdir/ic.o" "$dir"/{variables,ref_ic}.o \
        -L$MPI/lib -Wl,-rpath,$MPI/lib -lmpifort -lmpi -fuse-ld=lld
    echo "built $cdir/ref_ic"
  done < "$HERE/ref/ic_cases.txt"
}

# Multi-rank timing build: the reference's path at the benchmark grid with
# its own y-z decomposition (nyprocs x nzprocs MPI ranks) and the harness
# ref/ref_bench.f90.  Output: _ref/bench_<nx>_<npy>x<npz>/ref_bench.
build_bench() {
  local nx=$1 ny=$2 nz=$3 npy=$4 npz=$5
  local dir=$OUT/bench_${nx}_${npy}x${npz}
  mkdir -p "$dir"
  write_pin "$dir" $nx $ny $nz $npy $npz
  echo "      init_cond = cmplx(1.0_pr, 0.0_pr, pr)" > "$dir/ic.in"
  local flags="-O2 -flto -I$dir -I$MPI/include -module-dir $dir"
  for m in parameters constants error derivs ic variables solve; do
    $FC $flags -c "$REF/examples/ring/$m.f90" -o "$dir/$m.o"
  done
  $FC $flags -c "$HERE/ref/ref_bench.f90" -o "$dir/ref_bench.o"
  $FC -O2 -flto -o "$dir/ref_bench" "$dir"/{parameters,constants,error,derivs,ic,variables,solve,ref_bench}.o \
      -L$MPI/lib -Wl,-rpath,$MPI/lib -lmpifort -lmpi -flto -fuse-ld=lld
  echo "built $dir/ref_bench"
}

build_grid 64 64 64
build_grid 24 20 18
build_ics
build_bench 512 512 512 4 4
build_bench 32 32 32 2 2
build_bench 32 32 32 1 1
