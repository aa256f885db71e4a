! bromoc_gpu.f90 -- iso_c_binding front-end of libbromoc_gpu.so for the cg-gcmc-bd driver.
!
! Link this file's objects instead of energy.mono.o, dynamics0.mono.o, dynamics1.mono.o,
! dynamics2.mono.o, dynamics3.mono.o, dynamics0nuc.o, interact.o, grand.o, metropolis.o and
! repwalls.o:
! the externals at the end carry the reference routines' names and argument lists
! (energy.mono.f90:19, dynamics0.mono.f90:19, dynamics1.mono.f90:19, dynamics2.mono.f90:19,
! dynamics3.mono.f90:19, dynamics0nuc.f90:19, interact.f90:19, grand.f90:19,
! metropolis.f90:19, repwalls.f90:19), so simul1.f90 and shell_simul.f90 call the GPU path
! unchanged.  The first call hands the module state (grandmod, efpmod,
! extramod, nucleotmod, gsbpmod) to bromoc_create; positions travel to the device before
! every ENERGY/INTERACT and back after every DYNAMICS/GRAND.  A faster loop that keeps
! positions on the device for nbd steps is bromoc_gpu_bd(nbd) (see the caller notes at
! its definition).  Noise: the GPU path draws counter-based Philox Gaussians keyed by the
! run's seed instead of libgfortran rand(); GRAND takes its uniforms from rndm().
module bromoc_gpu
use iso_c_binding
implicit none

type, bind(C) :: bgpu_pair
  integer(c_int32_t) :: efpot, chr, lj, srpmf, nxi, nknots
  real(c_double) :: fct, epp4, sgp2, c0, c1, c2, c3, c4, dmi, dm2lo, dm2hi, sc1, sc2, sc3
end type

type, bind(C) :: bgpu_grid
  integer(c_int32_t) :: nclx, ncly, nclz, nmaps
  real(c_double) :: dcel, xbcen, ybcen, zbcen
  type(c_ptr) :: data
end type

! bromoc_analysis: the cycle statistics of simul1.f90 (rho, rDNA, g(r), prob2/prob3 and the
! virial pressure) gathered on the device; what = sum of 1 (rho), 2 (rDNA), 4 (g(r)),
! 8 (prob), 16 (pressure)
type, bind(C) :: bgpu_analysis
  integer(c_int32_t) :: what, nzmax
  real(c_double) :: zmini, idelz
  integer(c_int32_t) :: igr, npoints
  real(c_double) :: dr
  integer(c_int32_t) :: datom, reserved
  real(c_double) :: czmin, czmax, tvol, temp, ikbt
end type

type, bind(C) :: bgpu_system
  integer(c_int32_t) :: abi_version, ntot, nsites, nfix, capacity, ntype, nold, nttyp, npair
  integer(c_int32_t) :: flags
  type(c_ptr) :: x, y, z, typei, typtyp, namsite, stfree
  type(c_ptr) :: nwtype, cg, fact1a, fact2a, pairidx, pairs
  integer(c_int32_t) :: nknots, reserved0
  type(c_ptr) :: ep
  real(c_double) :: res, ires, rth, srpx, srpk, srpy
  real(c_double) :: cgnuc, epsnuc, scalepairing, fctn, kappa, ikappa, epsolv
  integer(c_int32_t) :: nbond, nangle, ndihe, nstack, nbp, nex, nqq, nsolv
  type(c_ptr) :: sitebond, distbond, siteangle, valangle, sitedihe, valdihe
  type(c_ptr) :: sitestack, sgstack, sitebp, sgbp, siteex, sgex, siteqq, siteslv
  type(bgpu_grid) :: phix, phiv
  real(c_double) :: svdw, thold8, vzmin, vzmax
  type(c_ptr) :: scal
  real(c_double) :: bdmax, fact0n1, fact0n2, lx2m, lx2p, ly2m, ly2p, lz2m, lz2p
  integer(c_int32_t) :: cntpts, reserved1
  type(c_ptr) :: zcont
  real(c_double) :: dt, kbtdt
  type(c_ptr) :: diffusion
  real(c_double) :: plength2, pcenter
  type(c_ptr) :: ampl3, p3
  integer(c_int32_t) :: nspline, reserved2
  type(c_ptr) :: xs, ys, bs, cs, ds
  type(bgpu_grid) :: rf
  type(c_ptr) :: gsrfen, greff
  real(c_double) :: sqrfac
  integer(c_int32_t) :: afn, ctn
  type(c_ptr) :: sn, af, csn, kx, ky, kz, contrx, contry, contrz
  integer(c_int32_t) :: unsplit, nsites1st, notrans_axes, reserved3
  real(c_double) :: notrx, notry, notrz, ibeta, diff0, diffcutoff, thold27
  integer(c_int32_t) :: qpore, reserved4
  real(c_double) :: voltage, afact, ceps, zmemb1, zmemb2
  type(c_ptr) :: ampl1, p1, ampl2, p2, rcylinder
  integer(c_int32_t) :: mmij_shape, xnpol, ynpol, znpol, ntpol, reserved5
  type(c_ptr) :: lstpx, lstpy, lstpz, bnorm, mmij
  real(c_double) :: xscal, yscal, zscal, rfscal
  real(c_double) :: cx, cy, cz, rsphe, rsphe2, iecx, iecy
end type

type, bind(C) :: bgpu_energies
  real(c_double) :: ener, eelec, evdw, esrpmf, esrpmfmx, eefpot, eefpotmx, eqqmx, evdwmx
  real(c_double) :: ebond, eang, edihe, estack, ebp, eex, eqq, esolv
  real(c_double) :: egsbpa, egsbpb, evdwgd, erfpar, econ, ememb
end type

type, bind(C) :: bgpu_gcmc
  integer(c_int32_t) :: nbuffer, reserved
  type(c_ptr) :: ibfftyp, avnum, mu, kb, lzmin, lzmax, bufferbias, ntotat
  real(c_double) :: lx, ly, cx, cy, cz, ikbt
  type(c_ptr) :: rmin, rmax
end type

! BROMOC_ABI_VERSION of include/bromoc_gpu.h: the layout of bgpu_system and bgpu_gcmc above
integer(c_int32_t), parameter :: BGPU_ABI_VERSION = 5
integer(c_int32_t), parameter :: F_PAR=1, F_NUCL=2, F_NONBOND=4, F_NOBOND=8, F_PHIX=16, &
     F_PHIV=32, F_SVDW=64, F_SOLV=128, F_DEBY=256, F_COUNTION=512, F_DNAFREE=1024, &
     F_MEMB=65536, F_MMIJ=131072, F_BSPLINE=262144, F_NMCDEN=524288, F_DEBYHYB=1048576, &
     F_PROXDIFF=2097152, F_NOTRANS=4194304, F_APFOR=8388608, F_CONTRANS=16777216, &
     F_RFPAR=33554432, F_SPHERE=67108864, F_ECYL=134217728, F_DIFFUSE=268435456, &
     F_PROFILE=536870912, F_RFPSIN=1073741824
integer(c_int32_t), parameter :: NOISE_PHILOX = 1

interface
  integer(c_int) function bromoc_create(s, device, ctx) bind(C, name='bromoc_create')
    import :: bgpu_system, c_int, c_ptr
    type(bgpu_system), intent(in) :: s
    integer(c_int), value :: device
    type(c_ptr), intent(out) :: ctx
  end function
  integer(c_int) function bromoc_set_positions(ctx, n, x, y, z, t) bind(C, name='bromoc_set_positions')
    import :: c_int, c_int32_t, c_double, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: n
    real(c_double), intent(in) :: x(*), y(*), z(*)
    integer(c_int32_t), intent(in) :: t(*)
  end function
  integer(c_int) function bromoc_get_positions(ctx, x, y, z, t, n) bind(C, name='bromoc_get_positions')
    import :: c_int, c_int32_t, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: x(*), y(*), z(*)
    integer(c_int32_t), intent(out) :: t(*), n
  end function
  integer(c_int) function bromoc_get_forces(ctx, fx, fy, fz) bind(C, name='bromoc_get_forces')
    import :: c_int, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), intent(out) :: fx(*), fy(*), fz(*)
  end function
  integer(c_int) function bromoc_energy(ctx, want, e) bind(C, name='bromoc_energy')
    import :: c_int, c_int32_t, c_ptr, bgpu_energies
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: want
    type(bgpu_energies), intent(out) :: e
  end function
  integer(c_int) function bromoc_step(ctx, n, mode, row, seed, ctr, e) bind(C, name='bromoc_step')
    import :: c_int, c_int32_t, c_int64_t, c_ptr, bgpu_energies
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: n, mode, row
    integer(c_int64_t), value :: seed, ctr
    type(bgpu_energies), intent(out) :: e
  end function
  integer(c_int) function bromoc_dynamics(ctx, mode, row, seed, ctr) bind(C, name='bromoc_dynamics')
    import :: c_int, c_int32_t, c_int64_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: mode, row
    integer(c_int64_t), value :: seed, ctr
  end function
  integer(c_int) function bromoc_get_flux(ctx, nf, nb) bind(C, name='bromoc_get_flux')
    import :: c_int, c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(out) :: nf(*), nb(*)
  end function
  integer(c_int) function bromoc_get_warn(ctx, w) bind(C, name='bromoc_get_warn')
    import :: c_int, c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(out) :: w(*)
  end function
  integer(c_int) function bromoc_interact(ctx, xj, yj, zj, jt, j, al, d) bind(C, name='bromoc_interact')
    import :: c_int, c_int32_t, c_double, c_ptr
    type(c_ptr), value :: ctx
    real(c_double), value :: xj, yj, zj
    integer(c_int32_t), value :: jt, j, al
    real(c_double), intent(out) :: d
  end function
  integer(c_int) function bromoc_set_gcmc(ctx, g) bind(C, name='bromoc_set_gcmc')
    import :: c_int, c_ptr, bgpu_gcmc
    type(c_ptr), value :: ctx
    type(bgpu_gcmc), intent(in) :: g
  end function
  integer(c_int) function bromoc_set_ibuffer(ctx, ib) bind(C, name='bromoc_set_ibuffer')
    import :: c_int, c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(in) :: ib(*)
  end function
  integer(c_int) function bromoc_get_ibuffer(ctx, ib) bind(C, name='bromoc_get_ibuffer')
    import :: c_int, c_int32_t, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), intent(out) :: ib(*)
  end function
  integer(c_int) function bromoc_grand(ctx, ng, u, nu, used, nat, prob, icyc, ener) bind(C, name='bromoc_grand')
    import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: ng
    real(c_double), intent(in) :: u(*)
    integer(c_int64_t), value :: nu
    integer(c_int64_t), intent(out) :: used
    integer(c_int32_t), intent(inout) :: nat(*), prob(*)
    integer(c_int64_t), value :: icyc
    real(c_double), intent(inout) :: ener
  end function
  integer(c_int) function bromoc_metropolis(ctx, nm, mcm, ikt, u, nu, used, done, ener) &
      bind(C, name='bromoc_metropolis')
    import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
    type(c_ptr), value :: ctx
    integer(c_int32_t), value :: nm
    real(c_double), value :: mcm, ikt
    real(c_double), intent(in) :: u(*)
    integer(c_int64_t), value :: nu
    integer(c_int64_t), intent(out) :: used
    integer(c_int32_t), intent(out) :: done
    real(c_double), intent(inout) :: ener
  end function
  integer(c_int) function bromoc_repwalls(nx, ny, nz, map, walls, device) bind(C, name='bromoc_repwalls')
    import :: c_int, c_int32_t, c_int8_t
    integer(c_int32_t), value :: nx, ny, nz, walls
    integer(c_int8_t), intent(inout) :: map(*)
    integer(c_int), value :: device
  end function
  integer(c_int) function bromoc_set_analysis(ctx, an) bind(C, name='bromoc_set_analysis')
    import :: c_int, c_ptr, bgpu_analysis
    type(c_ptr), value :: ctx
    type(bgpu_analysis), intent(in) :: an
  end function
  integer(c_int) function bromoc_analyze(ctx) bind(C, name='bromoc_analyze')
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function
  ! rho(ntype*nzmax), rdna(nttyp*nzmax), gr(ntype*npoints), prob2/prob3(ntype*(datom+1)):
  ! row-major with the type as the row, i.e. a Fortran array (nzmax, ntype) etc.; pass
  ! c_null_ptr for a statistic not gathered
  integer(c_int) function bromoc_get_analysis(ctx, rho, rdna, gr, irmax, prob2, prob3, pres, &
      nsteps) bind(C, name='bromoc_get_analysis')
    import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
    type(c_ptr), value :: ctx, rho, rdna, gr, prob2, prob3
    integer(c_int32_t), intent(out) :: irmax
    real(c_double), intent(out) :: pres
    integer(c_int64_t), intent(out) :: nsteps
  end function
  integer(c_int) function bromoc_destroy(ctx) bind(C, name='bromoc_destroy')
    import :: c_int, c_ptr
    type(c_ptr), value :: ctx
  end function
  type(c_ptr) function bromoc_last_error(ctx) bind(C, name='bromoc_last_error')
    import :: c_ptr
    type(c_ptr), value :: ctx
  end function
end interface

type(c_ptr), save :: gctx = c_null_ptr
logical, save :: sites_moved = .false.
integer(c_int64_t), save :: bd_counter = 0
integer(c_int64_t), save :: bd_seed = 20240613_c_int64_t
integer, save :: gcmc_ready = 0

contains

! Status check: the reference stops through error(...,faterr) (error.f90:30-32).
subroutine bgpu_check(rc, where)
  integer(c_int), intent(in) :: rc
  character(*), intent(in) :: where
  if (rc /= 0) then
    write(*,'(a,a,a,i0)') 'bromoc_gpu: ', where, ' failed, status ', rc
    stop 1
  endif
end subroutine

! Drop the device context: the next call hands the module state over again.  For a
! driver that re-reads its system (a new shell_simul input block).
subroutine bromoc_gpu_reset() bind(C, name='bromoc_gpu_reset')
  integer(c_int) :: rc
  if (c_associated(gctx)) rc = bromoc_destroy(gctx)
  gctx = c_null_ptr
  sites_moved = .false.
  gcmc_ready = 0
end subroutine

! Seed and step counter of the Philox noise (the reference seeds rand() from its input
! file's random seed, shell_simul.f90 iseed).
subroutine bromoc_gpu_set_seed(seed, counter) bind(C, name='bromoc_gpu_set_seed')
  integer(c_int64_t), value :: seed, counter
  bd_seed = seed
  bd_counter = counter
end subroutine

! Hand the reference's module state to the GPU library (once).
subroutine bromoc_gpu_init(seed)
  use grandmod
  use efpmod
  use extramod
  use nucleotmod
  use gsbpmod
  use apfmod
  use splinemod, only: sp_b => b, sp_c => c, sp_d => d, sp_x => xs, sp_y => ys, nspline
  integer(c_int64_t), intent(in), optional :: seed
  type(bgpu_system) :: s
  type(bgpu_pair), allocatable, target :: pr(:)
  integer(c_int32_t), allocatable, target :: pidx(:), ityp(:), tt(:), nm(:), nwt(:)
  integer(c_int32_t), allocatable, target :: lb(:), la(:), ld(:), ls(:), lp(:), le(:), lq(:), lv(:)
  integer(c_int8_t), allocatable, target :: sfree(:)
  real(c_double), allocatable, target :: xs(:), ys(:), zs(:), cgs(:), f1(:), f2(:), scs(:), zc(:)
  real(c_double), allocatable, target :: db(:), va(:), vd(:), ss(:), sb(:), se(:)
  real(c_double), allocatable, target :: epc(:,:)
  real(c_float), allocatable, target :: gx(:), rfs(:), rfr(:)
  real(c_double), allocatable, target :: dif(:), am3(:), pp3(:), spx(:), spy(:), spb(:), spc(:), spd(:)
  integer(c_int32_t), allocatable, target :: asn(:), ccsn(:)
  real(c_double), allocatable, target :: ma1(:), mp1(:), ma2(:), mp2(:), mrc(:)
  integer(c_int32_t), allocatable, target :: mlx(:), mly(:), mlz(:)
  real(c_double), allocatable, target :: mbn(:), mmm(:)
  real(c_double), allocatable, target :: aaf(:), ckx(:), cky(:), ckz(:), ccx(:), ccy(:), ccz(:)
  integer :: nc
  integer :: nsp, ngrid, nmaps
  integer(c_int8_t), allocatable, target :: gv(:)
  integer :: i, a, b, nk, nt1, ntc, nindex, cap
  integer(c_uint32_t) :: fl
  if (c_associated(gctx)) return
  if (present(seed)) bd_seed = seed
  fl = 0
  if (Qpar) fl = ior(fl, F_PAR)
  if (Qnucl) fl = ior(fl, F_NUCL)
  if (Qnonbond) fl = ior(fl, F_NONBOND)
  if (Qnobond) fl = ior(fl, F_NOBOND)
  if (Qphix) fl = ior(fl, F_PHIX)
  if (Qphiv) fl = ior(fl, F_PHIV)
  if (Qphiv .and. .not. Qtrln) fl = ior(fl, F_BSPLINE)
  if (Qsvdw) fl = ior(fl, F_SVDW)
  if (Qnucl .and. Qsolv) fl = ior(fl, F_SOLV)
  if (Qnucl .and. Qdeby) fl = ior(fl, F_DEBY)
  if (Qcountion) fl = ior(fl, F_COUNTION)
  if (Qnucl .and. Qdnafree) fl = ior(fl, F_DNAFREE)
  if (Qmemb) fl = ior(fl, F_MEMB)
  if (Qmmij) fl = ior(fl, F_MMIJ)
  if (Qnmcden) fl = ior(fl, F_NMCDEN)
  if (Qnucl .and. Qdebyhyb) fl = ior(fl, F_DEBYHYB)
  if (Qproxdiff) fl = ior(fl, F_PROXDIFF)
  if (Qnucl .and. Qnotrans) fl = ior(fl, F_NOTRANS)
  if (Qapfor) fl = ior(fl, F_APFOR)
  if (Qnucl .and. Qcontrans) fl = ior(fl, F_CONTRANS)
  if (Qrfpar) fl = ior(fl, F_RFPAR)
  if (Qrfpar .and. Qrfpsin) fl = ior(fl, F_RFPSIN)
  if (Qsphere) fl = ior(fl, F_SPHERE)
  if (Qecyl) fl = ior(fl, F_ECYL)
  if (Qdiffuse) fl = ior(fl, F_DIFFUSE)
  if (Qprofile) fl = ior(fl, F_PROFILE)
  cap = datom
  nt1 = nttyp + 1
  ntc = (nttyp + ndna + 1) * nion / 2
  allocate(xs(ntot), ys(ntot), zs(ntot), ityp(ntot))
  xs = x(1:ntot); ys = y(1:ntot); zs = z(1:ntot); ityp = typei(1:ntot)
  allocate(tt(max(1,nsites)), nm(max(1,nsites)), sfree(max(1,nsites)))
  do i = 1, nsites
    tt(i) = typtyp(i)
    select case (namsite(i))
    case ('S '); nm(i) = 0
    case ('P '); nm(i) = 1
    case ('Ab'); nm(i) = 2
    case ('Tb'); nm(i) = 3
    case ('Cb'); nm(i) = 4
    case default; nm(i) = 5
    end select
    sfree(i) = merge(1_c_int8_t, 0_c_int8_t, stfree(i))
  enddo
  allocate(nwt(ntype+1), cgs(ntype+1), f1(ntype+1), f2(ntype+1))
  nwt(1) = 0; cgs(1) = 0.0; f1(1) = 0.0; f2(1) = 0.0
  nwt(2:ntype+1) = nwtype(1:ntype); cgs(2:ntype+1) = cg(1:ntype)
  f1(2:ntype+1) = fact1a(1:ntype); f2(2:ntype+1) = fact2a(1:ntype)
  ! pair index table over atnam2 types: nindex.f90 where a pair is defined
  allocate(pidx(nt1*nt1))
  pidx = 0
  do a = 1, nttyp
    do b = 1, nttyp
      if (a > ndna .or. b > ndna) pidx(a*nt1 + b + 1) = nindex(a, b)
    enddo
  enddo
  allocate(pr(ntc+1))
  pr(1) = bgpu_pair(0,0,0,0,0,0, 0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0,0d0)
  do i = 1, ntc
    pr(i+1)%efpot = merge(1, 0, Qefpot(i)); pr(i+1)%chr = merge(1, 0, Qchr(i))
    pr(i+1)%lj = merge(1, 0, Qlj(i)); pr(i+1)%srpmf = merge(1, 0, Qsrpmfi(i))
    pr(i+1)%fct = fct(i); pr(i+1)%epp4 = epp4(i); pr(i+1)%sgp2 = sgp2(i)
    pr(i+1)%c0 = c0(i); pr(i+1)%c1 = c1(i); pr(i+1)%c2 = c2(i); pr(i+1)%c3 = c3(i); pr(i+1)%c4 = c4(i)
    pr(i+1)%nxi = 0; pr(i+1)%nknots = 0
    pr(i+1)%dmi = 0; pr(i+1)%dm2lo = 0; pr(i+1)%dm2hi = 0
    pr(i+1)%sc1 = 0; pr(i+1)%sc2 = 0; pr(i+1)%sc3 = 0
    if (Qefpot(i)) then
      pr(i+1)%nxi = nxi(i) - 1
      ! tables lie back to back in ep (shell_simul.f90:1826-1849): a table ends where the
      ! next one starts, the last one at the end of ep
      nk = size(ep, 2) + 1
      do b = 1, ntc
        if (Qefpot(b) .and. nxi(b) > nxi(i)) nk = min(nk, nxi(b))
      enddo
      pr(i+1)%nknots = nk - nxi(i)
      pr(i+1)%dmi = dmi(i); pr(i+1)%dm2lo = dm2(1,i); pr(i+1)%dm2hi = dm2(2,i)
      pr(i+1)%sc1 = sc(1,i); pr(i+1)%sc2 = sc(2,i); pr(i+1)%sc3 = sc(3,i)
    endif
  enddo
  allocate(epc(3, max(1, size(ep, 2))))
  epc = 0.0
  if (allocated(ep)) epc(:, 1:size(ep,2)) = ep
  ! DNA lists, component-major and 0-based
  call pack2(sitebond, nbond, lb); call pack2(sitestack, nstack, ls); call pack2(sitebp, nbp, lp)
  call pack2(siteex, nex, le); call pack2(siteqq, nqq, lq); call pack2(siteslv, nsolv, lv)
  allocate(la(max(1,3*nangle)), ld(max(1,4*ndihe)))
  do i = 1, 3
    if (nangle > 0) la((i-1)*nangle+1:i*nangle) = siteangle(1:nangle, i) - 1
  enddo
  do i = 1, 4
    if (ndihe > 0) ld((i-1)*ndihe+1:i*ndihe) = sitedihe(1:ndihe, i) - 1
  enddo
  allocate(db(max(1,nbond)), va(max(1,nangle)), vd(max(1,ndihe)), ss(max(1,nstack)), &
           sb(max(1,nbp)), se(max(1,nex)))
  if (nbond > 0) db(1:nbond) = distbond(1:nbond)
  if (nangle > 0) va(1:nangle) = valangle(1:nangle)
  if (ndihe > 0) vd(1:ndihe) = valdihe(1:ndihe)
  if (nstack > 0) ss(1:nstack) = sgstack(1:nstack)
  if (nbp > 0) sb(1:nbp) = sgbp(1:nbp)
  if (nex > 0) se(1:nex) = sgex(1:nex)
  allocate(scs(nt1), zc(max(1,cntpts)))
  scs = 1.0
  if (Qphiv .and. Qsvdw) scs(2:nt1) = scal(1:nttyp)
  if (cntpts > 0) zc(1:cntpts) = zcont(1:cntpts)

  s%abi_version = BGPU_ABI_VERSION
  s%ntot = ntot; s%nsites = nsites; s%nfix = nfix; s%capacity = cap
  s%ntype = ntype; s%nold = nold; s%nttyp = nttyp; s%npair = ntc; s%flags = int(fl, c_int32_t)
  s%x = c_loc(xs); s%y = c_loc(ys); s%z = c_loc(zs); s%typei = c_loc(ityp)
  s%typtyp = c_loc(tt); s%namsite = c_loc(nm); s%stfree = c_loc(sfree)
  s%nwtype = c_loc(nwt); s%cg = c_loc(cgs); s%fact1a = c_loc(f1); s%fact2a = c_loc(f2)
  s%pairidx = c_loc(pidx); s%pairs = c_loc(pr)
  s%nknots = size(epc, 2); s%ep = c_loc(epc)
  s%res = res; s%ires = ires; s%rth = rth; s%srpx = srpx; s%srpk = srpk; s%srpy = srpy
  s%cgnuc = cgnuc; s%epsnuc = epsnuc; s%scalepairing = scalepairing; s%fctn = fctn
  s%kappa = kappa; s%ikappa = ikappa; s%epsolv = epsolv
  s%nbond = nbond; s%nangle = nangle; s%ndihe = ndihe; s%nstack = nstack; s%nbp = nbp
  s%nex = nex; s%nqq = nqq; s%nsolv = nsolv
  s%sitebond = c_loc(lb); s%distbond = c_loc(db); s%siteangle = c_loc(la); s%valangle = c_loc(va)
  s%sitedihe = c_loc(ld); s%valdihe = c_loc(vd); s%sitestack = c_loc(ls); s%sgstack = c_loc(ss)
  s%sitebp = c_loc(lp); s%sgbp = c_loc(sb); s%siteex = c_loc(le); s%sgex = c_loc(se)
  s%siteqq = c_loc(lq); s%siteslv = c_loc(lv)
  s%phix = bgpu_grid(0, 0, 0, 0, 0d0, 0d0, 0d0, 0d0, c_null_ptr)
  s%phiv = bgpu_grid(0, 0, 0, 0, 0d0, 0d0, 0d0, 0d0, c_null_ptr)
  if (Qphix) then
    allocate(gx(size(phix)))
    gx = phix
    s%phix = bgpu_grid(nclx1, ncly1, nclz1, 0, dcel1, xbcen1, ybcen1, zbcen1, c_loc(gx))
  endif
  if (Qphiv) then
    ! with NMCDEN the library reads the six site maps (with DNA) and one map per ion type,
    ! contiguous; readphi.f90:188-192 allocates phiv over the last ion map only when
    ! PARTICLE is on, so a shorter phiv is refused rather than read past its end
    nmaps = 1
    if (Qnmcden) then
      nmaps = 0
      if (Qnucl) nmaps = 6
      if (Qpar) nmaps = nmaps + ntype - nold
    endif
    if (size(phiv, kind=8) /= int(nmaps, 8) * nclx2 * ncly2 * nclz2) &
      call bgpu_check(-1_c_int, 'phiv size (NMCDEN maps)')
    allocate(gv(size(phiv)))
    gv = phiv
    s%phiv = bgpu_grid(nclx2, ncly2, nclz2, nmaps, dcel2, xbcen2, ybcen2, zbcen2, c_loc(gv))
  endif
  s%svdw = svdw; s%thold8 = thold8; s%vzmin = vzmin; s%vzmax = vzmax; s%scal = c_loc(scs)
  s%bdmax = bdmax; s%fact0n1 = fact0n1; s%fact0n2 = fact0n2
  s%lx2m = lx2m; s%lx2p = lx2p; s%ly2m = ly2m; s%ly2p = ly2p; s%lz2m = lz2m; s%lz2p = lz2p
  s%cntpts = merge(cntpts, 0, Qcountion); s%zcont = c_loc(zc)
  ! position-dependent diffusion (simul1.f90:342; DIFFUSION and PROFILE commands)
  allocate(dif(ntype+1), am3(ntype+1), pp3(ntype+1))
  dif(1) = 0.0; am3(1) = 0.0; pp3(1) = 1.0
  dif(2:) = diffusion(1:ntype); am3(2:) = ampl3(1:ntype); pp3(2:) = p3(1:ntype)
  s%dt = dt; s%kbtdt = kbtdt; s%diffusion = c_loc(dif)
  s%plength2 = plength2; s%pcenter = pcenter; s%ampl3 = c_loc(am3); s%p3 = c_loc(pp3)
  nsp = merge(nspline, 0, Qprofile)
  allocate(spx(max(1,nsp)), spy(max(1,nsp)), spb(max(1,nsp)), spc(max(1,nsp)), spd(max(1,nsp)))
  if (nsp > 0) then
    spx = sp_x(1:nsp); spy = sp_y(1:nsp); spb = sp_b(1:nsp); spc = sp_c(1:nsp); spd = sp_d(1:nsp)
  endif
  s%nspline = nsp
  s%xs = c_loc(spx); s%ys = c_loc(spy); s%bs = c_loc(spb); s%cs = c_loc(spc); s%ds = c_loc(spd)
  ! reaction-field parameter grids (gsbpmod, rfpar.f90:430-445)
  s%rf = bgpu_grid(0, 0, 0, 0, 0d0, 0d0, 0d0, 0d0, c_null_ptr)
  if (Qrfpar) then
    s%rf = bgpu_grid(nclx3, ncly3, nclz3, 0, dcel3, xbcen3, ybcen3, zbcen3, c_null_ptr)
    ngrid = size(gsrfen)
    allocate(rfs(ngrid), rfr(ngrid))
    rfs = gsrfen; rfr = greff
  else
    allocate(rfs(1), rfr(1))
    rfs = 0.0; rfr = 0.0
  endif
  s%gsrfen = c_loc(rfs); s%greff = c_loc(rfr); s%sqrfac = sqrfac
  ! APFOR (apfmod) and CONTRANS / NOTRANS (nucleotmod), shell_simul.f90:650-770
  s%afn = 0; s%ctn = 0
  allocate(asn(1), aaf(3))
  asn = 0; aaf = 0.0
  if (Qapfor .and. afn > 0) then
    deallocate(asn, aaf)
    allocate(asn(afn), aaf(3*afn))
    asn = sn(1:afn) - 1
    aaf = reshape(af(1:3, 1:afn), [3*afn])
    s%afn = afn
  endif
  s%sn = c_loc(asn); s%af = c_loc(aaf)
  nc = 0
  if (Qnucl .and. Qcontrans) nc = ctn
  allocate(ccsn(max(1,nc)), ckx(max(1,nc)), cky(max(1,nc)), ckz(max(1,nc)), ccx(nc+1), &
           ccy(nc+1), ccz(nc+1))
  ccsn = 0; ckx = 0.0; cky = 0.0; ckz = 0.0; ccx = 0.0; ccy = 0.0; ccz = 0.0
  if (nc > 0) then
    ccsn(1:nc) = csn(1:nc); ckx(1:nc) = kx(1:nc); cky(1:nc) = ky(1:nc); ckz(1:nc) = kz(1:nc)
    ccx = contrx(1:nc+1); ccy = contry(1:nc+1); ccz = contrz(1:nc+1)
  endif
  s%ctn = nc; s%csn = c_loc(ccsn); s%kx = c_loc(ckx); s%ky = c_loc(cky); s%kz = c_loc(ckz)
  s%contrx = c_loc(ccx); s%contry = c_loc(ccy); s%contrz = c_loc(ccz)
  s%unsplit = merge(1, 0, Qnucl .and. Qcontrans .and. Qunsplit); s%nsites1st = nsites1st
  s%notrans_axes = 0
  if (Qnucl .and. Qnotrans) then
    if (Qnotrx) s%notrans_axes = ior(s%notrans_axes, 1)
    if (Qnotry) s%notrans_axes = ior(s%notrans_axes, 2)
    if (Qnotrz) s%notrans_axes = ior(s%notrans_axes, 4)
  endif
  s%notrx = notrx; s%notry = notry; s%notrz = notrz
  ! PROXDIFF (grandmod beta, diff0, diffcutoff) and the B-spline wall threshold
  s%ibeta = ibeta; s%diff0 = diff0; s%diffcutoff = diffcutoff
  s%thold27 = thold27
  ! MEMBRANE (grandmod, shell_simul.f90:1563-1640), per atnam2 type
  allocate(ma1(nt1), mp1(2*nt1), ma2(nt1), mp2(2*nt1), mrc(nt1))
  ma1 = 0.0; mp1 = 1.0; ma2 = 0.0; mp2 = 1.0; mrc = 1.0
  if (Qmemb) then
    ma1(2:nt1) = ampl1(1:nttyp); ma2(2:nt1) = ampl2(1:nttyp); mrc(2:nt1) = rcylinder(1:nttyp)
    do i = 1, nttyp
      mp1(2*i+1) = p1(1,i); mp1(2*i+2) = p1(2,i)
      mp2(2*i+1) = p2(1,i); mp2(2*i+2) = p2(2,i)
    enddo
  endif
  s%qpore = merge(1, 0, Qmemb .and. Qpore)
  s%voltage = voltage; s%afact = afact; s%ceps = ceps; s%zmemb1 = zmemb1; s%zmemb2 = zmemb2
  s%ampl1 = c_loc(ma1); s%p1 = c_loc(mp1); s%ampl2 = c_loc(ma2); s%p2 = c_loc(mp2)
  s%rcylinder = c_loc(mrc)
  ! MMIJ reaction field (gsbpmod, readmmij.f90 + rect_norm.f90 as shell_simul.f90:2780-2831
  ! leaves it); shape 2 (SPHERE) is refused by bromoc_create
  allocate(mlx(max(1,ntpol)), mly(max(1,ntpol)), mlz(max(1,ntpol)), mbn(max(1,ntpol)))
  allocate(mmm(max(1,ntpol*ntpol)))
  mlx = 0; mly = 0; mlz = 0; mbn = 0.0; mmm = 0.0
  s%mmij_shape = 0; s%ntpol = 0
  if (Qmmij) then
    s%mmij_shape = merge(1, 2, shapes == 'RECTBOX ')
    s%ntpol = ntpol
    if (shapes == 'RECTBOX ') then
      s%xnpol = xnpol; s%ynpol = ynpol; s%znpol = znpol
      mlx(1:ntpol) = lstpx(1:ntpol); mly(1:ntpol) = lstpy(1:ntpol); mlz(1:ntpol) = lstpz(1:ntpol)
    endif
    mbn(1:ntpol) = bnorm(1:ntpol)
    mmm(1:ntpol*ntpol) = mmij(1:ntpol*ntpol)
    s%xscal = xscal; s%yscal = yscal; s%zscal = zscal; s%rfscal = rfscal
  endif
  s%lstpx = c_loc(mlx); s%lstpy = c_loc(mly); s%lstpz = c_loc(mlz)
  s%bnorm = c_loc(mbn); s%mmij = c_loc(mmm)
  s%cx = cx; s%cy = cy; s%cz = cz
  s%rsphe = rsphe; s%rsphe2 = rsphe2; s%iecx = iecx; s%iecy = iecy
  call bgpu_check(bromoc_create(s, 0, gctx), 'bromoc_create')
end subroutine

subroutine pack2(l, n, out)
  integer, intent(in) :: l(:,:)
  integer, intent(in) :: n
  integer(c_int32_t), allocatable, intent(out) :: out(:)
  allocate(out(max(1, 2*n)))
  if (n > 0) then
    out(1:n) = l(1:n, 1) - 1
    out(n+1:2*n) = l(1:n, 2) - 1
  endif
end subroutine

subroutine bromoc_gpu_upload()
  use grandmod
  integer(c_int32_t), allocatable :: t(:)
  allocate(t(ntot))
  t = typei(1:ntot)
  call bgpu_check(bromoc_set_positions(gctx, int(ntot, c_int32_t), x, y, z, t), 'bromoc_set_positions')
end subroutine

subroutine bromoc_gpu_download()
  use grandmod
  integer(c_int32_t) :: n
  integer(c_int32_t), allocatable :: t(:)
  allocate(t(datom))
  call bgpu_check(bromoc_get_positions(gctx, x, y, z, t, n), 'bromoc_get_positions')
  ntot = n
  typei(1:ntot) = t(1:ntot)
end subroutine

subroutine bromoc_gpu_counters()
  use grandmod
  use nucleotmod
  integer(c_int32_t), allocatable :: nf(:), nb(:), w(:)
  integer :: i, j, nt
  nt = ntype - nold
  if (Qcountion .and. cntpts > 0) then
    allocate(nf(nt*cntpts), nb(nt*cntpts))
    call bgpu_check(bromoc_get_flux(gctx, nf, nb), 'bromoc_get_flux')
    do i = 1, nt
      do j = 1, cntpts
        nforward(i, j) = nf((i-1)*cntpts + j)
        nbackward(i, j) = nb((i-1)*cntpts + j)
      enddo
    enddo
  endif
  allocate(w(nttyp+1))
  call bgpu_check(bromoc_get_warn(gctx, w), 'bromoc_get_warn')
  warn(1:nttyp) = w(2:nttyp+1)
end subroutine

! nbd BD steps (ENERGY + DYNAMICS0/1/2/3 + DYNAMICS0NUC + NOTRANS) with the positions
! kept on the device: the replacement for simul1.f90:363-390 when the pressure virial is off.
subroutine bromoc_gpu_bd(nbd)
  use grandmod
  integer, intent(in) :: nbd
  type(bgpu_energies) :: e
  call bromoc_gpu_init()
  call bromoc_gpu_upload()
  call bgpu_check(bromoc_step(gctx, int(nbd, c_int32_t), NOISE_PHILOX, 0_c_int32_t, bd_seed, &
                              bd_counter, e), 'bromoc_step')
  bd_counter = bd_counter + nbd
  call bromoc_gpu_download()
  call bgpu_check(bromoc_get_forces(gctx, fx, fy, fz), 'bromoc_get_forces')
  call bromoc_gpu_counters()
  call bgpu_store_energies(e)
end subroutine

subroutine bgpu_store_energies(e)
  use grandmod
  use gsbpmod
  use nucleotmod
  type(bgpu_energies), intent(in) :: e
  ener = e%ener; eelec = e%eelec; evdw = e%evdw; esrpmf = e%esrpmf; esrpmfmx = e%esrpmfmx
  eefpot = e%eefpot; eefpotmx = e%eefpotmx; eqqmx = e%eqqmx; evdwmx = e%evdwmx
  ebond = e%ebond; eang = e%eang; edihe = e%edihe; estack = e%estack; ebp = e%ebp
  eex = e%eex; eqq = e%eqq; esolv = e%esolv; egsbpa = e%egsbpa; egsbpb = e%egsbpb
  evdwgd = e%evdwgd; erfpar = e%erfpar; econ = e%econ; ememb = e%ememb
end subroutine

end module bromoc_gpu

! ------------------------------------------------------------ drop-in externals
SUBROUTINE ENERGY
use bromoc_gpu
use grandmod
implicit none
type(bgpu_energies) :: e
call bromoc_gpu_init()
call bromoc_gpu_upload()
call bgpu_check(bromoc_energy(gctx, merge(1_c_int32_t, 0_c_int32_t, Qforces), e), 'bromoc_energy')
if (Qforces) call bgpu_check(bromoc_get_forces(gctx, fx, fy, fz), 'bromoc_get_forces')
call bromoc_gpu_counters()
call bgpu_store_energies(e)
sites_moved = .false.
end subroutine

! The device pass moves every moving ion (and every free site): the ranges the driver
! hands DYNAMICS0..3 and DYNAMICS0NUC (simul1.f90:369-381); any other range is refused
! rather than ignored.
subroutine bgpu_check_range(ninit, nfinal, lo, hi, where)
  integer, intent(in) :: ninit, nfinal, lo, hi
  character(*), intent(in) :: where
  if (ninit /= lo .or. nfinal /= hi) then
    write(*,'(a,a,a,i0,a,i0,a,i0,a,i0)') 'bromoc_gpu: ', where, ' moves ', lo, '..', hi, &
         ' only, not ', ninit, '..', nfinal
    stop 1
  endif
end subroutine

! DYNAMICS0 moves the ions and, in the same device pass, the free DNA sites; the
! DYNAMICS0NUC that follows it in simul1.f90:380 then has nothing left to do.
SUBROUTINE DYNAMICS0(ninit, nfinal)
use bromoc_gpu
use grandmod
use nucleotmod
implicit none
integer ninit, nfinal
call bgpu_check_range(ninit, nfinal, nsites + nfix + 1, ntot, 'DYNAMICS')
call bromoc_gpu_init()
call bgpu_check(bromoc_dynamics(gctx, NOISE_PHILOX, 0_c_int32_t, bd_seed, bd_counter), 'bromoc_dynamics')
bd_counter = bd_counter + 1
call bromoc_gpu_download()
call bromoc_gpu_counters()
sites_moved = .true.
end subroutine

! DYNAMICS1 (switch diffusion) and DYNAMICS2 (spline profile): the device pass picks the
! ion update from the flags handed over at set-up, as simul1.f90:370-378 does.
SUBROUTINE DYNAMICS1(ninit, nfinal)
implicit none
integer ninit, nfinal
call DYNAMICS0(ninit, nfinal)
end subroutine

SUBROUTINE DYNAMICS2(ninit, nfinal)
implicit none
integer ninit, nfinal
call DYNAMICS0(ninit, nfinal)
end subroutine

! DYNAMICS3 (DNA-proximity diffusion): the device pass recomputes each ion's nearest-site
! record from the same positions as the preceding ENERGY.
SUBROUTINE DYNAMICS3(ninit, nfinal)
implicit none
integer ninit, nfinal
call DYNAMICS0(ninit, nfinal)
end subroutine

SUBROUTINE DYNAMICS0NUC(ninit, nfinal)
use bromoc_gpu
use nucleotmod
implicit none
integer ninit, nfinal
call bgpu_check_range(ninit, nfinal, 1, nsites, 'DYNAMICS0NUC')
if (sites_moved) then
  sites_moved = .false.
  return
endif
call bromoc_gpu_init()
call bgpu_check(bromoc_dynamics(gctx, NOISE_PHILOX, 0_c_int32_t, bd_seed, bd_counter), 'bromoc_dynamics')
bd_counter = bd_counter + 1
call bromoc_gpu_download()
end subroutine

SUBROUTINE INTERACT(dener, xj, yj, zj, jtype, j, Qalert)
use bromoc_gpu
implicit none
real dener, xj, yj, zj
integer jtype, j
logical*1 Qalert
real(c_double) :: d
call bromoc_gpu_init()
call bromoc_gpu_upload()
call bgpu_check(bromoc_interact(gctx, xj, yj, zj, int(jtype, c_int32_t), int(j - 1, c_int32_t), &
                                merge(1_c_int32_t, 0_c_int32_t, Qalert), d), 'bromoc_interact')
dener = d
end subroutine

! GRAND draws every uniform from the reference's rndm() up front (at most 6 per attempt in
! a box; in a sphere or elliptic cylinder INSERT redraws its 3 until the point is inside, so
! 768 more per attempt are drawn) and hands them to the device loop, which consumes them in
! grand.f90's order.
SUBROUTINE GRAND(ngcmc, prob, icycle)
use bromoc_gpu
use grandmod
implicit none
integer ngcmc, prob(nbuffer, 0:datom)
integer*8 icycle
real(c_double), allocatable :: u(:)
integer(c_int32_t), allocatable :: natv(:), pr(:), ib(:)
integer(c_int64_t) :: used
type(bgpu_gcmc) :: g
integer(c_int32_t), allocatable, target :: bt(:), tot(:)
real(c_double), allocatable, target :: av(:), muv(:), kbv(:), zl(:), zh(:), rl(:), rh(:)
integer(c_int8_t), allocatable, target :: bb(:)
real(c_double) :: e
integer :: k, b, n
call bromoc_gpu_init()
call bromoc_gpu_upload()
if (gcmc_ready == 0) then
  allocate(bt(nbuffer+1), tot(nbuffer+1), av(nbuffer+1), muv(nbuffer+1), kbv(nbuffer+1), &
           zl(nbuffer+1), zh(nbuffer+1), bb(nbuffer+1), rl(nbuffer+1), rh(nbuffer+1))
  bt(1) = 0; tot(1) = 0; av(1) = 0; muv(1) = 0; kbv(1) = 0; zl(1) = 0; zh(1) = 0; bb(1) = 0
  rl(1) = 0; rh(1) = 0
  rl(2:) = Rmin(1:nbuffer); rh(2:) = Rmax(1:nbuffer)
  bt(2:) = ibfftyp(1:nbuffer); tot(2:) = ntotat(1:nbuffer); av(2:) = avnum(1:nbuffer)
  muv(2:) = mu(1:nbuffer); kbv(2:) = kb(1:nbuffer); zl(2:) = LZmin(1:nbuffer); zh(2:) = LZmax(1:nbuffer)
  do b = 1, nbuffer
    bb(b+1) = merge(1_c_int8_t, 0_c_int8_t, Qbufferbias(b))
  enddo
  g = bgpu_gcmc(nbuffer, 0, c_loc(bt), c_loc(av), c_loc(muv), c_loc(kbv), c_loc(zl), c_loc(zh), &
                c_loc(bb), c_loc(tot), lx, ly, cx, cy, cz, ikbt, c_loc(rl), c_loc(rh))
  call bgpu_check(bromoc_set_gcmc(gctx, g), 'bromoc_set_gcmc')
  gcmc_ready = 1
endif
allocate(ib(ntot))
ib = ibuffer(1:ntot)
call bgpu_check(bromoc_set_ibuffer(gctx, ib), 'bromoc_set_ibuffer')
if (Qsphere .or. Qecyl) then
  allocate(u(774*ngcmc + 1))
else
  allocate(u(6*ngcmc + 1))
endif
do k = 1, size(u)
  u(k) = rndm()
enddo
allocate(natv(nbuffer+1), pr(nbuffer*(datom+1)))
natv(1) = 0
natv(2:) = nat(1:nbuffer)
do b = 1, nbuffer
  do n = 0, datom
    pr((b-1)*(datom+1) + n + 1) = prob(b, n)
  enddo
enddo
e = ener
call bgpu_check(bromoc_grand(gctx, int(ngcmc, c_int32_t), u, int(size(u), c_int64_t), used, natv, pr, &
                             icycle, e), 'bromoc_grand')
ener = e
nat(1:nbuffer) = natv(2:)
do b = 1, nbuffer
  do n = 0, datom
    prob(b, n) = pr((b-1)*(datom+1) + n + 1)
  enddo
enddo
call bromoc_gpu_download()
deallocate(ib)
allocate(ib(ntot))
call bgpu_check(bromoc_get_ibuffer(gctx, ib), 'bromoc_get_ibuffer')
ibuffer(1:ntot) = ib
end subroutine

! METROPOLIS draws its uniforms from the reference's rndm() ahead of the device loop: 5 per
! trial and a margin; trials that the batch cannot finish (the box-rejection loop of MOVE
! draws 4 more per retry) run in the next call with the unused uniforms first.
SUBROUTINE METROPOLIS(nmcm)
use bromoc_gpu
use grandmod
implicit none
integer nmcm
real(c_double), allocatable :: u(:), w(:)
integer(c_int64_t) :: used
integer(c_int32_t) :: done, left
real(c_double) :: e
integer :: k, nu, have
call bromoc_gpu_init()
call bromoc_gpu_upload()
left = nmcm
have = 0
nu = 5*nmcm + 64
allocate(u(nu))
e = ener
do while (left > 0)
  do k = have + 1, nu
    u(k) = rndm()
  enddo
  call bgpu_check(bromoc_metropolis(gctx, left, mcmax, ikbt, u, int(nu, c_int64_t), used, done, e), &
                  'bromoc_metropolis')
  left = left - done
  have = nu - int(used)
  u(1:have) = u(int(used)+1:nu)
  if (done == 0) then          ! one trial needs more than the batch holds: widen it
    allocate(w(2*nu))
    w(1:have) = u(1:have)
    call move_alloc(w, u)
    nu = 2*nu
  endif
enddo
ener = e
call bromoc_gpu_download()
end subroutine

! REPWALLS layers the repulsion map (gsbpmod phiv) on the device, one stencil pass a layer.
SUBROUTINE REPWALLS(walls)
use bromoc_gpu
use gsbpmod
implicit none
integer*1 walls
call bgpu_check(bromoc_repwalls(int(nclx2, c_int32_t), int(ncly2, c_int32_t), &
                int(nclz2, c_int32_t), phiv, int(walls, c_int32_t), 0_c_int), 'bromoc_repwalls')
end subroutine
