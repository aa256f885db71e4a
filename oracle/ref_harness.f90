! ref_harness.f90 -- drives the reference cg-gcmc-bd routines (compiled unchanged from
! their own sources by oracle/Makefile) from a bromoc_system description.
!
! TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg as the
! reference checker; never part of the product path.
!
! What this file supplies, and why:
!  * ref_* entry points (bind(C)) that copy a bromoc_system into the reference's module
!    globals (grandmod, efpmod, extramod, nucleotmod, gsbpmod, apfmod) and call ENERGY,
!    DYNAMICS0, DYNAMICS0NUC, INTERACT, GRAND and SPLINEPOT themselves.
!  * rgauss(), rndm(), rndm1(): the reference draws noise from gfortran's rand()
!    (random.f90:19-31, rgauss.f90:19-26).  Here they return injected values in call
!    order, the seam through which GPU and reference consume identical noise.
module ref_harness
use iso_c_binding
implicit none

type, bind(C) :: c_pair
  integer(c_int32_t) :: efpot, chr, lj, srpmf, nxi, nknots
  real(c_double) :: fct, epp4, sgp2, c0, c1, c2, c3, c4, dmi, dm2lo, dm2hi, sc1, sc2, sc3
end type

type, bind(C) :: c_grid
  integer(c_int32_t) :: nclx, ncly, nclz, reserved
  real(c_double) :: dcel, xbcen, ybcen, zbcen
  type(c_ptr) :: data
end type

type, bind(C) :: c_system
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
  type(c_grid) :: phix, phiv
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
  type(c_grid) :: rf
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

type, bind(C) :: c_gcmc
  integer(c_int32_t) :: nbuffer, reserved
  type(c_ptr) :: ibfftyp, avnum, mu, kb, lzmin, lzmax, bufferbias, ntotat
  real(c_double) :: lx, ly, cx, cy, cz, ikbt
  type(c_ptr) :: rmin, rmax
end type

! injected streams
real(c_double), allocatable :: gstream(:), ustream(:)
integer :: gpos = 0, upos = 0, ulen = 0
integer(c_int32_t), parameter :: F_PAR=1, F_NUCL=2, F_NONBOND=4, F_NOBOND=8, F_PHIX=16, F_PHIV=32, &
     F_SVDW=64, F_SOLV=128, F_DEBY=256, F_COUNTION=512, F_DNAFREE=1024, &
     F_RFPAR=33554432, F_SPHERE=67108864, F_ECYL=134217728, F_DIFFUSE=268435456, &
     F_PROFILE=536870912, F_RFPSIN=1073741824, &
     F_BSPLINE=262144, F_PROXDIFF=2097152, F_NOTRANS=4194304, F_APFOR=8388608, &
     F_CONTRANS=16777216, F_MEMB=65536, F_MMIJ=131072, F_DEBYHYB=1048576, &
     F_NMCDEN=524288
integer :: site_names_code(0:5)

contains

logical(1) function flag(f, b)
  integer(c_int32_t), intent(in) :: f, b
  flag = iand(f, b) /= 0
end function

end module ref_harness

! ---------------------------------------------------------------- noise seams
function rgauss()
use ref_harness
implicit none
real rgauss
gpos = gpos + 1
rgauss = gstream(gpos)
end function

function rndm()
use ref_harness
implicit none
real rndm
upos = upos + 1
if (upos > ulen) then
  rndm = 0.5
else
  rndm = ustream(upos)
endif
end function

function rndm1()
use ref_harness
implicit none
real rndm1
upos = upos + 1
if (upos > ulen) then
  rndm1 = 0.5
else
  rndm1 = 1.0 - ustream(upos)
endif
end function

! ------------------------------------------------------------------ set-up
subroutine ref_init(s) bind(C, name='ref_init')
use ref_harness
use grandmod
use efpmod
use extramod
use nucleotmod
use gsbpmod
use apfmod
implicit none
type(c_system), intent(in) :: s
real(c_double), pointer :: xp(:), yp(:), zp(:), dp(:)
integer(c_int32_t), pointer :: ip(:)
integer(c_int8_t), pointer :: bp(:)
real(c_float), pointer :: fp(:)
type(c_pair), pointer :: pp(:)
integer :: i, n, ns, ntc, ncel3, nt, nmaps

character(len=2), parameter :: names(0:5) = (/'S ','P ','Ab','Tb','Cb','Gb'/)

ntot = s%ntot
nsites = s%nsites
nfix = s%nfix
ntype = s%ntype
nold = s%nold
nttyp = s%nttyp
natom = ntot - nfix              ! shell_simul.f90:2435
Qpar = flag(s%flags, F_PAR)
Qnucl = flag(s%flags, F_NUCL)
Qnonbond = flag(s%flags, F_NONBOND)
Qnobond = flag(s%flags, F_NOBOND)
Qphix = flag(s%flags, F_PHIX)
Qphiv = flag(s%flags, F_PHIV)
Qsvdw = flag(s%flags, F_SVDW)
Qsolv = flag(s%flags, F_SOLV)
Qdeby = flag(s%flags, F_DEBY)
Qcountion = flag(s%flags, F_COUNTION)
Qdnafree = flag(s%flags, F_DNAFREE)
Qenergy = .true.
Qforces = .false.
Qmemb = flag(s%flags, F_MEMB)
Qmmij = flag(s%flags, F_MMIJ)
Qrfpar = .false.
Qtrln = .not. flag(s%flags, F_BSPLINE)
Qnmcden = flag(s%flags, F_NMCDEN)
! the map numbering of vdwgd1trln.f90:74-76 counts ion types from inuc+1 (one strand):
! nold nucleotide types (shell_simul.f90:530)
istrs = 1
inuc = s%nold
Qproxdiff = flag(s%flags, F_PROXDIFF)
Qdebyhyb = flag(s%flags, F_DEBYHYB)
Qapfor = flag(s%flags, F_APFOR)
Qcontrans = flag(s%flags, F_CONTRANS)
Qwarn = .false.
Qsphere = flag(s%flags, F_SPHERE)
Qecyl = flag(s%flags, F_ECYL)
cx = s%cx; cy = s%cy; cz = s%cz
rsphe = s%rsphe; rsphe2 = s%rsphe2; iecx = s%iecx; iecy = s%iecy
Qnotrans = flag(s%flags, F_NOTRANS)
Qdiffuse = flag(s%flags, F_DIFFUSE)
Qprofile = flag(s%flags, F_PROFILE)
Qrfpar = flag(s%flags, F_RFPAR)
Qrfpsin = flag(s%flags, F_RFPSIN)

call c_f_pointer(s%x, xp, [ntot]); x(1:ntot) = xp
call c_f_pointer(s%y, yp, [ntot]); y(1:ntot) = yp
call c_f_pointer(s%z, zp, [ntot]); z(1:ntot) = zp
call c_f_pointer(s%typei, ip, [ntot]); typei(1:ntot) = ip
fx = 0.0; fy = 0.0; fz = 0.0
ns = nsites
if (ns > 0) then
  call c_f_pointer(s%typtyp, ip, [ns]); typtyp(1:ns) = ip
  if (allocated(namsite)) deallocate(namsite)
  if (allocated(stfree)) deallocate(stfree)
  allocate(namsite(ns), stfree(ns))
  call c_f_pointer(s%namsite, ip, [ns])
  do i = 1, ns
    namsite(i) = names(ip(i))
  enddo
  call c_f_pointer(s%stfree, bp, [ns])
  do i = 1, ns
    stfree(i) = bp(i) /= 0
  enddo
endif
call c_f_pointer(s%nwtype, ip, [ntype+1]); nwtype(1:ntype) = ip(2:ntype+1)
call c_f_pointer(s%cg, dp, [ntype+1]); cg(1:ntype) = dp(2:ntype+1)
call c_f_pointer(s%fact1a, dp, [ntype+1]); fact1a(1:ntype) = dp(2:ntype+1)
call c_f_pointer(s%fact2a, dp, [ntype+1]); fact2a(1:ntype) = dp(2:ntype+1)

! pair tables: the harness's nindex is the host table pairidx (nindex.f90 semantics)
ntc = s%npair
if (allocated(Qefpot)) deallocate(Qefpot, Qchr, Qlj, Qsrpmfi, fct, epp4, sgp2, c0, c1, c2, c3, c4, &
                                  nxi, dmi, dm2, sc)
allocate(Qefpot(ntc), Qchr(ntc), Qlj(ntc), Qsrpmfi(ntc), fct(ntc), epp4(ntc), sgp2(ntc), &
         c0(ntc), c1(ntc), c2(ntc), c3(ntc), c4(ntc), nxi(ntc), dmi(ntc), dm2(2,ntc), sc(3,ntc))
call c_f_pointer(s%pairs, pp, [ntc+1])
do i = 1, ntc
  Qefpot(i) = pp(i+1)%efpot /= 0
  Qchr(i) = pp(i+1)%chr /= 0
  Qlj(i) = pp(i+1)%lj /= 0
  Qsrpmfi(i) = pp(i+1)%srpmf /= 0
  fct(i) = pp(i+1)%fct
  epp4(i) = pp(i+1)%epp4
  sgp2(i) = pp(i+1)%sgp2
  c0(i) = pp(i+1)%c0; c1(i) = pp(i+1)%c1; c2(i) = pp(i+1)%c2; c3(i) = pp(i+1)%c3; c4(i) = pp(i+1)%c4
  nxi(i) = pp(i+1)%nxi + 1
  dmi(i) = pp(i+1)%dmi
  dm2(1,i) = pp(i+1)%dm2lo
  dm2(2,i) = pp(i+1)%dm2hi
  sc(1,i) = pp(i+1)%sc1; sc(2,i) = pp(i+1)%sc2; sc(3,i) = pp(i+1)%sc3
enddo
call set_pairidx(s)
if (allocated(ep)) deallocate(ep)
allocate(ep(3, max(1, s%nknots)))
if (s%nknots > 0) then
  call c_f_pointer(s%ep, dp, [3*s%nknots])
  ep = reshape(dp, [3, s%nknots])
endif
res = s%res
ires = s%ires
rth = s%rth; srpx = s%srpx; srpk = s%srpk; srpy = s%srpy

! CG-DNA
cgnuc = s%cgnuc; epsnuc = s%epsnuc; scalepairing = s%scalepairing; fctn = s%fctn
kappa = s%kappa; ikappa = s%ikappa; epsolv = s%epsolv
call set_list2(s%sitebond, s%distbond, s%nbond, 1)
call set_list3(s)
call set_list4(s)
call set_list2(s%sitestack, s%sgstack, s%nstack, 2)
call set_list2(s%sitebp, s%sgbp, s%nbp, 3)
call set_list2(s%siteex, s%sgex, s%nex, 4)
call set_list2(s%siteqq, c_null_ptr, s%nqq, 5)
call set_list2(s%siteslv, c_null_ptr, s%nsolv, 6)

! grids (readphi.f90:49-51,160-170,330-340)
if (Qphix) then
  nclx1 = s%phix%nclx; ncly1 = s%phix%ncly; nclz1 = s%phix%nclz
  dcel1 = s%phix%dcel; idcel1 = 1.0/dcel1
  tranx1 = 0.5*(nclx1-1)*dcel1; trany1 = 0.5*(ncly1-1)*dcel1; tranz1 = 0.5*(nclz1-1)*dcel1
  xbcen1 = s%phix%xbcen; ybcen1 = s%phix%ybcen; zbcen1 = s%phix%zbcen
  ncel3 = nclx1*ncly1*nclz1
  if (allocated(phix)) deallocate(phix)
  allocate(phix(ncel3))
  call c_f_pointer(s%phix%data, fp, [ncel3]); phix = fp
endif
if (Qphiv) then
  nclx2 = s%phiv%nclx; ncly2 = s%phiv%ncly; nclz2 = s%phiv%nclz
  dcel2 = s%phiv%dcel; idcel2 = 1.0/dcel2
  tranx2 = 0.5*(nclx2-1)*dcel2; trany2 = 0.5*(ncly2-1)*dcel2; tranz2 = 0.5*(nclz2-1)*dcel2
  xbcen2 = s%phiv%xbcen; ybcen2 = s%phiv%ybcen; zbcen2 = s%phiv%zbcen
  ncel3 = nclx2*ncly2*nclz2
  nmaps = 1
  if (Qnmcden) then   ! readphi.f90:176-240: six site maps, then one per ion type
    nmaps = 0
    if (Qnucl) nmaps = 6
    if (Qpar) nmaps = nmaps + ntype - nold
  endif
  if (allocated(phiv)) deallocate(phiv)
  allocate(phiv(ncel3*nmaps))
  call c_f_pointer(s%phiv%data, bp, [ncel3*nmaps]); phiv = bp
  svdw = s%svdw; thold8 = s%thold8; vzmin = s%vzmin; vzmax = s%vzmax; thold27 = s%thold27
  if (allocated(scal)) deallocate(scal)
  allocate(scal(nttyp))
  if (Qsvdw) then
    call c_f_pointer(s%scal, dp, [nttyp+1]); scal(1:nttyp) = dp(2:nttyp+1)
  else
    scal = 1.0
  endif
endif
if (allocated(warn)) deallocate(warn)
allocate(warn(nttyp))
warn = 0

! position-dependent diffusion (DIFFUSION / PROFILE commands, simul1.f90:342)
dt = s%dt
kbtdt = s%kbtdt
if (Qdiffuse .or. Qprofile .or. Qproxdiff) then
  call c_f_pointer(s%diffusion, dp, [ntype+1]); diffusion(1:ntype) = dp(2:ntype+1)
endif
! PROXDIFF command (shell_simul.f90:1903-1910)
if (Qproxdiff) then
  ibeta = s%ibeta; diff0 = s%diff0; diffcutoff = s%diffcutoff
endif
dids = 0.0
call set_dna_options(s)
if (Qmemb) call set_membrane(s)
if (Qmmij) call set_mmij(s)
if (Qdiffuse) then
  plength2 = s%plength2; pcenter = s%pcenter
  call c_f_pointer(s%ampl3, dp, [ntype+1]); ampl3(1:ntype) = dp(2:ntype+1)
  call c_f_pointer(s%p3, dp, [ntype+1]); p3(1:ntype) = dp(2:ntype+1)
endif
if (Qprofile) call set_profile(s)
! reaction-field parameter grids (rfpar.f90:430-445)
if (Qrfpar) then
  nclx3 = s%rf%nclx; ncly3 = s%rf%ncly; nclz3 = s%rf%nclz
  dcel3 = s%rf%dcel; idcel3 = 1.0/dcel3
  tranx3 = 0.5*(nclx3-1)*dcel3; trany3 = 0.5*(ncly3-1)*dcel3; tranz3 = 0.5*(nclz3-1)*dcel3
  xbcen3 = s%rf%xbcen; ybcen3 = s%rf%ybcen; zbcen3 = s%rf%zbcen
  ncel3 = nclx3*ncly3*nclz3
  if (Qrfpsin) then
    n = ncel3
  else
    n = ncel3*(ntype - nold)
  endif
  if (allocated(gsrfen)) deallocate(gsrfen, greff)
  allocate(gsrfen(n), greff(n))
  call c_f_pointer(s%gsrfen, fp, [n]); gsrfen = fp
  call c_f_pointer(s%greff, fp, [n]); greff = fp
  sqrfac = s%sqrfac
endif

! dynamics (simul1.f90:342-349 precomputed by the caller)
bdmax = s%bdmax; fact0n1 = s%fact0n1; fact0n2 = s%fact0n2
lx2m = s%lx2m; lx2p = s%lx2p; ly2m = s%ly2m; ly2p = s%ly2p; lz2m = s%lz2m; lz2p = s%lz2p
cntpts = s%cntpts
nt = max(1, ntype - nold)
if (allocated(nforward)) deallocate(nforward, nbackward)
allocate(nforward(nt, max(1,cntpts)), nbackward(nt, max(1,cntpts)))
nforward = 0; nbackward = 0
if (allocated(zcont)) deallocate(zcont)
allocate(zcont(max(1,cntpts)))
if (cntpts > 0) then
  call c_f_pointer(s%zcont, dp, [cntpts]); zcont(1:cntpts) = dp
endif
n = 0
end subroutine

subroutine set_profile(s)
use ref_harness
use splinemod, only: sp_b => b, sp_c => c, sp_d => d, xs, ys, nspline
implicit none
type(c_system), intent(in) :: s
real(c_double), pointer :: dp(:)
integer :: n
n = s%nspline
nspline = n
call c_f_pointer(s%xs, dp, [n]); xs(1:n) = dp
call c_f_pointer(s%ys, dp, [n]); ys(1:n) = dp
call c_f_pointer(s%bs, dp, [n]); sp_b(1:n) = dp
call c_f_pointer(s%cs, dp, [n]); sp_c(1:n) = dp
call c_f_pointer(s%ds, dp, [n]); sp_d(1:n) = dp
end subroutine

! MMIJ state (gsbpmod) as shell_simul.f90:2780-2831, READMMIJ and RECT_NORM leave it for
! a RECTBOX region: the region centred at 0, degrees, norms, the symmetric matrix
subroutine set_mmij(s)
use ref_harness
use gsbpmod
implicit none
type(c_system), intent(in) :: s
integer(c_int32_t), pointer :: ip(:)
real(c_double), pointer :: dp(:)
integer :: n
n = s%ntpol
shapes = 'RECTBOX '
xnpol = s%xnpol; ynpol = s%ynpol; znpol = s%znpol; ntpol = n
if (allocated(lstpx)) deallocate(lstpx, lstpy, lstpz, bnorm, coef, mmij)
allocate(lstpx(n), lstpy(n), lstpz(n), bnorm(n), coef(n), mmij(n*n))
call c_f_pointer(s%lstpx, ip, [n]); lstpx = ip
call c_f_pointer(s%lstpy, ip, [n]); lstpy = ip
call c_f_pointer(s%lstpz, ip, [n]); lstpz = ip
call c_f_pointer(s%bnorm, dp, [n]); bnorm = dp
call c_f_pointer(s%mmij, dp, [n*n]); mmij = dp
xscal = s%xscal; yscal = s%yscal; zscal = s%zscal; rfscal = s%rfscal
xmin = -1.0/xscal; xmax = 1.0/xscal
ymin = -1.0/yscal; ymax = 1.0/yscal
zmin = -1.0/zscal; zmax = 1.0/zscal
end subroutine

! MEMBRANE state as shell_simul.f90:1563-1640 leaves it (per atnam2 type)
subroutine set_membrane(s)
use ref_harness
use grandmod
implicit none
type(c_system), intent(in) :: s
real(c_double), pointer :: dp(:)
integer :: n
n = s%nttyp
Qpore = s%qpore /= 0
voltage = s%voltage; afact = s%afact; ceps = s%ceps; zmemb1 = s%zmemb1; zmemb2 = s%zmemb2
call c_f_pointer(s%ampl1, dp, [n+1]); ampl1(1:n) = dp(2:n+1)
call c_f_pointer(s%ampl2, dp, [n+1]); ampl2(1:n) = dp(2:n+1)
call c_f_pointer(s%rcylinder, dp, [n+1]); rcylinder(1:n) = dp(2:n+1)
call c_f_pointer(s%p1, dp, [2*(n+1)]); p1(1:2, 1:n) = reshape(dp(3:2*n+2), [2, n])
call c_f_pointer(s%p2, dp, [2*(n+1)]); p2(1:2, 1:n) = reshape(dp(3:2*n+2), [2, n])
end subroutine

! APFOR (apfmod), CONTRANS and NOTRANS (nucleotmod) state as shell_simul.f90:650-770 leaves it
subroutine set_dna_options(s)
use ref_harness
use apfmod
use nucleotmod
implicit none
type(c_system), intent(in) :: s
integer(c_int32_t), pointer :: ip(:)
real(c_double), pointer :: dp(:)
integer :: n
if (nsites > 0) insites = 1.0/nsites
if (Qapfor) then
  n = s%afn
  afn = n
  if (allocated(sn)) deallocate(sn, af)
  allocate(sn(max(1,n)), af(3, max(1,n)))
  if (n > 0) then
    call c_f_pointer(s%sn, ip, [n]); sn(1:n) = ip + 1
    call c_f_pointer(s%af, dp, [3*n]); af(:, 1:n) = reshape(dp, [3, n])
  endif
endif
if (Qcontrans) then
  n = s%ctn
  ctn = n
  if (allocated(kx)) deallocate(kx, ky, kz, csn, contrx, contry, contrz)
  allocate(kx(max(1,n)), ky(max(1,n)), kz(max(1,n)), csn(max(1,n)), contrx(n+1), contry(n+1), &
           contrz(n+1))
  if (n > 0) then
    call c_f_pointer(s%csn, ip, [n]); csn(1:n) = ip
    call c_f_pointer(s%kx, dp, [n]); kx(1:n) = dp
    call c_f_pointer(s%ky, dp, [n]); ky(1:n) = dp
    call c_f_pointer(s%kz, dp, [n]); kz(1:n) = dp
  endif
  call c_f_pointer(s%contrx, dp, [n+1]); contrx(1:n+1) = dp
  call c_f_pointer(s%contry, dp, [n+1]); contry(1:n+1) = dp
  call c_f_pointer(s%contrz, dp, [n+1]); contrz(1:n+1) = dp
  Qunsplit = s%unsplit /= 0
  nsites1st = s%nsites1st
endif
if (Qnotrans) then
  Qnotrx = iand(s%notrans_axes, 1) /= 0
  Qnotry = iand(s%notrans_axes, 2) /= 0
  Qnotrz = iand(s%notrans_axes, 4) /= 0
  notrx = s%notrx; notry = s%notry; notrz = s%notrz
endif
end subroutine

! SPLINE (spline.f90:19-98) on caller arrays
subroutine ref_spline(n, x, y, b, c, d) bind(C, name='ref_spline')
use iso_c_binding
implicit none
integer(c_int32_t), value :: n
real(c_double), intent(in) :: x(n), y(n)
real(c_double), intent(out) :: b(n), c(n), d(n)
real, allocatable :: xx(:), yy(:), bb(:), cc(:), dd(:)
integer :: m
m = n
allocate(xx(m), yy(m), bb(m), cc(m), dd(m))
xx = x; yy = y
call spline(m, xx, yy, bb, cc, dd)
b = bb; c = cc; d = dd
end subroutine

subroutine set_list2(lp, vp, n, which)
use ref_harness
use nucleotmod
implicit none
type(c_ptr), intent(in) :: lp, vp
integer(c_int32_t), intent(in) :: n
integer, intent(in) :: which
integer(c_int32_t), pointer :: ip(:)
real(c_double), pointer :: dp(:)
integer :: m
m = max(1, n)
if (n > 0) call c_f_pointer(lp, ip, [2*n])
if (n > 0 .and. c_associated(vp)) call c_f_pointer(vp, dp, [n])
select case (which)
case (1)
  nbond = n
  if (allocated(sitebond)) deallocate(sitebond, distbond)
  allocate(sitebond(m,2), distbond(m))
  if (n > 0) then
    sitebond(1:n,1) = ip(1:n) + 1; sitebond(1:n,2) = ip(n+1:2*n) + 1; distbond(1:n) = dp
  endif
case (2)
  nstack = n
  if (allocated(sitestack)) deallocate(sitestack, sgstack)
  allocate(sitestack(m,2), sgstack(m))
  if (n > 0) then
    sitestack(1:n,1) = ip(1:n) + 1; sitestack(1:n,2) = ip(n+1:2*n) + 1; sgstack(1:n) = dp
  endif
case (3)
  nbp = n
  if (allocated(sitebp)) deallocate(sitebp, sgbp)
  allocate(sitebp(m,2), sgbp(m))
  if (n > 0) then
    sitebp(1:n,1) = ip(1:n) + 1; sitebp(1:n,2) = ip(n+1:2*n) + 1; sgbp(1:n) = dp
  endif
case (4)
  nex = n
  if (allocated(siteex)) deallocate(siteex, sgex)
  allocate(siteex(m,2), sgex(m))
  if (n > 0) then
    siteex(1:n,1) = ip(1:n) + 1; siteex(1:n,2) = ip(n+1:2*n) + 1; sgex(1:n) = dp
  endif
case (5)
  nqq = n
  if (allocated(siteqq)) deallocate(siteqq)
  allocate(siteqq(m,2))
  if (n > 0) then
    siteqq(1:n,1) = ip(1:n) + 1; siteqq(1:n,2) = ip(n+1:2*n) + 1
  endif
case (6)
  nsolv = n
  if (allocated(siteslv)) deallocate(siteslv)
  allocate(siteslv(m,2))
  if (n > 0) then
    siteslv(1:n,1) = ip(1:n) + 1; siteslv(1:n,2) = ip(n+1:2*n) + 1
  endif
end select
end subroutine

subroutine set_list3(s)
use ref_harness
use nucleotmod
implicit none
type(c_system), intent(in) :: s
integer(c_int32_t), pointer :: ip(:)
real(c_double), pointer :: dp(:)
integer :: n, k
n = s%nangle
nangle = n
if (allocated(siteangle)) deallocate(siteangle, valangle)
allocate(siteangle(max(1,n),3), valangle(max(1,n)))
if (n > 0) then
  call c_f_pointer(s%siteangle, ip, [3*n]); call c_f_pointer(s%valangle, dp, [n])
  do k = 1, 3
    siteangle(1:n,k) = ip((k-1)*n+1:k*n) + 1
  enddo
  valangle(1:n) = dp
endif
end subroutine

subroutine set_list4(s)
use ref_harness
use nucleotmod
implicit none
type(c_system), intent(in) :: s
integer(c_int32_t), pointer :: ip(:)
real(c_double), pointer :: dp(:)
integer :: n, k
n = s%ndihe
ndihe = n
if (allocated(sitedihe)) deallocate(sitedihe, valdihe)
allocate(sitedihe(max(1,n),4), valdihe(max(1,n)))
if (n > 0) then
  call c_f_pointer(s%sitedihe, ip, [4*n]); call c_f_pointer(s%valdihe, dp, [n])
  do k = 1, 4
    sitedihe(1:n,k) = ip((k-1)*n+1:k*n) + 1
  enddo
  valdihe(1:n) = dp
endif
end subroutine

! nindex() of the reference maps atnam2 pairs through ndna/nion bookkeeping set up by
! reasig.f90; this harness sets that bookkeeping so nindex reproduces the host's pairidx,
! and aborts if it cannot.
subroutine set_pairidx(s)
use ref_harness
use nucleotmod
implicit none
type(c_system), intent(in) :: s
integer(c_int32_t), pointer :: ip(:)
integer :: a, b, n, nindex
n = s%nttyp
call c_f_pointer(s%pairidx, ip, [(n+1)*(n+1)])
do a = 1, n
  do b = 1, n
    if (ip(a*(n+1)+b+1) /= 0) then
      if (nindex(a, b) /= ip(a*(n+1)+b+1)) then
        write(*,*) 'ref_harness: pairidx does not follow nindex.f90 for ', a, b, &
                   nindex(a, b), ip(a*(n+1)+b+1)
        stop 3
      endif
    endif
  enddo
enddo
end subroutine

subroutine ref_set_nindex(ndna_in, nion_in) bind(C, name='ref_set_nindex')
use iso_c_binding
use nucleotmod
implicit none
integer(c_int32_t), value :: ndna_in, nion_in
ndna = ndna_in
nion = nion_in
ndnaxnion = ndna*nion
end subroutine

! --------------------------------------------------------------- ENERGY
subroutine ref_energy(qf, e) bind(C, name='ref_energy')
use ref_harness
use grandmod
use gsbpmod
use nucleotmod
implicit none
integer(c_int32_t), value :: qf
real(c_double), intent(out) :: e(23)
Qforces = qf /= 0
! terms ENERGY sets only under their option (erfpar under Qrfpar, energy.mono.f90:110-114)
! start from zero, not from a previous system's value
erfpar = 0.0; egsbpb = 0.0; egsbpa = 0.0; evdwgd = 0.0; ememb = 0.0; econ = 0.0
call energy
e = 0.0
e(1) = ener
e(2) = eelec; e(3) = evdw; e(4) = esrpmf; e(5) = esrpmfmx; e(6) = eefpot; e(7) = eefpotmx
e(8) = eqqmx; e(9) = evdwmx; e(10) = ebond; e(11) = eang; e(12) = edihe; e(13) = estack
e(14) = ebp; e(15) = eex; e(16) = eqq; e(17) = esolv; e(18) = egsbpa; e(19) = egsbpb
e(20) = evdwgd; e(21) = erfpar; e(22) = econ; e(23) = ememb
end subroutine

subroutine ref_get_forces(ofx, ofy, ofz) bind(C, name='ref_get_forces')
use ref_harness
use grandmod
implicit none
real(c_double), intent(out) :: ofx(*), ofy(*), ofz(*)
ofx(1:ntot) = fx(1:ntot); ofy(1:ntot) = fy(1:ntot); ofz(1:ntot) = fz(1:ntot)
end subroutine

subroutine ref_get_positions(ox, oy, oz, ot, n) bind(C, name='ref_get_positions')
use ref_harness
use grandmod
implicit none
real(c_double), intent(out) :: ox(*), oy(*), oz(*)
integer(c_int32_t), intent(out) :: ot(*), n
n = ntot
ox(1:ntot) = x(1:ntot); oy(1:ntot) = y(1:ntot); oz(1:ntot) = z(1:ntot)
ot(1:ntot) = typei(1:ntot)
end subroutine

subroutine ref_set_positions(n, ix, iy, iz, it) bind(C, name='ref_set_positions')
use ref_harness
use grandmod
implicit none
integer(c_int32_t), value :: n
real(c_double), intent(in) :: ix(*), iy(*), iz(*)
integer(c_int32_t), intent(in) :: it(*)
ntot = n
natom = n
x(1:n) = ix(1:n); y(1:n) = iy(1:n); z(1:n) = iz(1:n); typei(1:n) = it(1:n)
end subroutine

subroutine ref_get_flux(nf, nb) bind(C, name='ref_get_flux')
use ref_harness
use grandmod
implicit none
integer(c_int32_t), intent(out) :: nf(*), nb(*)
integer :: i, j, nt
nt = ntype - nold
do i = 1, nt
  do j = 1, cntpts
    nf((i-1)*cntpts + j) = nforward(i, j)
    nb((i-1)*cntpts + j) = nbackward(i, j)
  enddo
enddo
end subroutine

subroutine ref_get_warn(w) bind(C, name='ref_get_warn')
use ref_harness
use grandmod
use nucleotmod
implicit none
integer(c_int32_t), intent(out) :: w(*)
w(1) = 0
w(2:nttyp+1) = warn(1:nttyp)
end subroutine

! ------------------------------------------------------------- DYNAMICS
! One BD update with the forces left by ENERGY: DYNAMICS0 over moving ions then
! DYNAMICS0NUC over sites (simul1.f90:369-381).  g is [ntot][3] per particle; the
! stream handed to rgauss() follows the reference's draw order.
subroutine ref_dynamics(g) bind(C, name='ref_dynamics')
use ref_harness
use grandmod
use nucleotmod
implicit none
real(c_double), intent(in) :: g(3, *)
integer :: i, k
if (allocated(gstream)) deallocate(gstream)
allocate(gstream(3*ntot + 3))
k = 0
if (Qpar) then
  do i = nsites+nfix+1, ntot
    gstream(k+1:k+3) = g(1:3, i); k = k + 3
  enddo
endif
if (Qnucl) then
  do i = 1, nsites
    if (stfree(i)) then
      gstream(k+1:k+3) = g(1:3, i); k = k + 3
    endif
  enddo
endif
gpos = 0
! simul1.f90:369-382
if (Qpar) then
  if (Qdiffuse) then
    call dynamics1(nsites+nfix+1, ntot)
  elseif (Qprofile) then
    call dynamics2(nsites+nfix+1, ntot)
  elseif (Qproxdiff) then
    call dynamics3(nsites+nfix+1, ntot)
  else
    call dynamics0(nsites+nfix+1, ntot)
  endif
endif
if (Qnucl) call dynamics0nuc(1, nsites)
end subroutine

subroutine ref_get_dids(o) bind(C, name='ref_get_dids')
use ref_harness
use grandmod
implicit none
real(c_double), intent(out) :: o(5, *)
o(:, 1:ntot) = dids(:, 1:ntot)
end subroutine

! the NOTRANS lines of the driver's step loop, simul1.f90:382-386
subroutine ref_notrans() bind(C, name='ref_notrans')
use ref_harness
use grandmod
use nucleotmod
implicit none
if (Qnucl .and. Qnotrans) then
  if (Qnotrx) x(1:nsites)=x(1:nsites)-sum(x(1:nsites))*insites+notrx
  if (Qnotry) y(1:nsites)=y(1:nsites)-sum(y(1:nsites))*insites+notry
  if (Qnotrz) z(1:nsites)=z(1:nsites)-sum(z(1:nsites))*insites+notrz
endif
end subroutine

! METROPOLIS (metropolis.f90:19-58) with injected uniforms; natom follows ntot
subroutine ref_metropolis(nmcm, mcmax_in, ikbt_in, u, nu, used, ener_io) &
    bind(C, name='ref_metropolis')
use ref_harness
use grandmod
implicit none
integer(c_int32_t), value :: nmcm
real(c_double), value :: mcmax_in, ikbt_in
integer(c_int64_t), value :: nu
real(c_double), intent(in) :: u(*)
integer(c_int64_t), intent(out) :: used
real(c_double), intent(inout) :: ener_io
integer :: n
if (allocated(ustream)) deallocate(ustream)
allocate(ustream(max(1_8, nu)))
ustream(1:nu) = u(1:nu)
ulen = int(nu)
upos = 0
mcmax = mcmax_in
ikbt = ikbt_in
natom = ntot - nfix
ener = ener_io
n = nmcm
call metropolis(n)
ener_io = ener
used = upos
end subroutine

! ------------------------------------------------------------- INTERACT
subroutine ref_interact(xj, yj, zj, jtype, j, alert, dener) bind(C, name='ref_interact')
use ref_harness
implicit none
real(c_double), value :: xj, yj, zj
integer(c_int32_t), value :: jtype, j, alert
real(c_double), intent(out) :: dener
real :: d, xx, yy, zz
integer :: jt, jj
logical(1) :: qa
xx = xj; yy = yj; zz = zj; jt = jtype; jj = j + 1
qa = alert /= 0
call interact(d, xx, yy, zz, jt, jj, qa)
dener = d
end subroutine

! ---------------------------------------------------------------- GRAND
subroutine ref_set_gcmc(g, ib_in) bind(C, name='ref_set_gcmc')
use ref_harness
use grandmod
implicit none
type(c_gcmc), intent(in) :: g
integer(c_int32_t), intent(in) :: ib_in(*)
integer(c_int32_t), pointer :: ip(:)
real(c_double), pointer :: dp(:)
integer(c_int8_t), pointer :: bp(:)
integer :: nb, i
nb = g%nbuffer
nbuffer = nb
call c_f_pointer(g%ibfftyp, ip, [nb+1]); ibfftyp(1:nb) = ip(2:nb+1)
call c_f_pointer(g%ntotat, ip, [nb+1]); ntotat(1:nb) = ip(2:nb+1)
call c_f_pointer(g%avnum, dp, [nb+1]); avnum(1:nb) = dp(2:nb+1)
call c_f_pointer(g%mu, dp, [nb+1]); mu(1:nb) = dp(2:nb+1)
call c_f_pointer(g%kb, dp, [nb+1]); kb(1:nb) = dp(2:nb+1)
call c_f_pointer(g%lzmin, dp, [nb+1]); LZmin(1:nb) = dp(2:nb+1)
call c_f_pointer(g%lzmax, dp, [nb+1]); LZmax(1:nb) = dp(2:nb+1)
call c_f_pointer(g%bufferbias, bp, [nb+1])
do i = 1, nb
  Qbufferbias(i) = bp(i+1) /= 0
enddo
lx = g%lx; ly = g%ly; cx = g%cx; cy = g%cy; cz = g%cz; ikbt = g%ikbt
if (c_associated(g%rmin)) then
  call c_f_pointer(g%rmin, dp, [nb+1]); Rmin(1:nb) = dp(2:nb+1)
  call c_f_pointer(g%rmax, dp, [nb+1]); Rmax(1:nb) = dp(2:nb+1)
endif
ibuffer(1:ntot) = ib_in(1:ntot)
ninsert = 0; nremove = 0
end subroutine

subroutine ref_get_ibuffer(ib_out) bind(C, name='ref_get_ibuffer')
use ref_harness
use grandmod
implicit none
integer(c_int32_t), intent(out) :: ib_out(*)
ib_out(1:ntot) = ibuffer(1:ntot)
end subroutine

subroutine ref_grand(ngcmc, u, nu, used, nat_io, prob_io, pstride, icycle, ener_io) &
    bind(C, name='ref_grand')
use ref_harness
use grandmod
implicit none
integer(c_int32_t), value :: ngcmc, pstride
integer(c_int64_t), value :: nu, icycle
real(c_double), intent(in) :: u(*)
integer(c_int64_t), intent(out) :: used
integer(c_int32_t), intent(inout) :: nat_io(*), prob_io(*)
real(c_double), intent(inout) :: ener_io
integer, allocatable :: prob(:,:)
integer :: ib, n, ng
integer(8) :: ic
if (allocated(ustream)) deallocate(ustream)
allocate(ustream(max(1_8, nu)))
ustream(1:nu) = u(1:nu)
ulen = int(nu)
upos = 0
nat(1:nbuffer) = nat_io(2:nbuffer+1)
allocate(prob(nbuffer, 0:datom))
prob = 0
do ib = 1, nbuffer
  do n = 0, pstride-1
    prob(ib, n) = prob_io((ib-1)*pstride + n + 1)
  enddo
enddo
ener = ener_io
ng = ngcmc
ic = icycle
call grand(ng, prob, ic)
ener_io = ener
nat_io(2:nbuffer+1) = nat(1:nbuffer)
do ib = 1, nbuffer
  do n = 0, pstride-1
    prob_io((ib-1)*pstride + n + 1) = prob(ib, n)
  enddo
enddo
used = upos
deallocate(prob)
end subroutine

! ---------------------------------------------------------- table build
! SPLINEPOT (splinepot.f90:19-96) for one pair type into a scratch slot.
subroutine ref_build_table(n, xx, yy, chr, fctv, ep_out, pout) bind(C, name='ref_build_table')
use ref_harness
use efpmod
use extramod
implicit none
integer(c_int32_t), value :: n, chr
real(c_double), value :: fctv
real(c_double), intent(in) :: xx(*), yy(*)
real(c_double), intent(out) :: ep_out(3, *)
type(c_pair), intent(inout) :: pout
real, allocatable :: xs(:), ys(:)
if (allocated(ep)) deallocate(ep)
if (allocated(nxi)) deallocate(nxi, dmi, dm2, sc)
if (allocated(fct)) deallocate(fct, Qchr)
allocate(ep(3, n), nxi(1), dmi(1), dm2(2,1), sc(3,1), fct(1), Qchr(1))
allocate(xs(n), ys(n))
xs = xx(1:n); ys = yy(1:n)
nxi(1) = 1
fct(1) = fctv
Qchr(1) = chr /= 0
call splinepot(1, n, xs, ys, n)
ep_out(1:3, 1:n) = ep(1:3, 1:n)
pout%dmi = dmi(1); pout%dm2lo = dm2(1,1); pout%dm2hi = dm2(2,1)
pout%sc1 = sc(1,1); pout%sc2 = sc(2,1); pout%sc3 = sc(3,1)
pout%nknots = n; pout%efpot = 1; pout%chr = chr; pout%fct = fctv
deallocate(xs, ys)
end subroutine

! DISCRETIZE (discretize.f90:19-120) of one pair type from its analytic parameters into a
! scratch slot: samples at k*res for k = nnp..mnp
subroutine ref_discretize(pin, nnp, mnp, resv, rthv, srpxv, srpkv, srpyv, ep_out, pout) &
    bind(C, name='ref_discretize')
use ref_harness
use efpmod
use extramod
use grandmod
implicit none
type(c_pair), intent(in) :: pin
integer(c_int32_t), value :: nnp, mnp
real(c_double), value :: resv, rthv, srpxv, srpkv, srpyv
real(c_double), intent(out) :: ep_out(3, *)
type(c_pair), intent(inout) :: pout
integer :: n, nxf1
n = mnp - nnp + 1
if (allocated(ep)) deallocate(ep)
if (allocated(nxi)) deallocate(nxi, dmi, dm2, sc)
if (allocated(fct)) deallocate(fct, Qchr)
if (allocated(epp4)) deallocate(epp4, sgp2)
if (allocated(c0)) deallocate(c0, c1, c2, c3, c4)
if (allocated(Qlj)) deallocate(Qlj, Qsrpmfi, Qefpot)
allocate(ep(3, n), nxi(1), dmi(1), dm2(2,1), sc(3,1), fct(1), Qchr(1))
allocate(epp4(1), sgp2(1), c0(1), c1(1), c2(1), c3(1), c4(1), Qlj(1), Qsrpmfi(1), Qefpot(1))
nxi(1) = 1
nxf1 = n
res = resv; ires = 1.0/resv
rth = rthv; srpx = srpxv; srpk = srpkv; srpy = srpyv
fct(1) = pin%fct; Qchr(1) = pin%chr /= 0
epp4(1) = pin%epp4; sgp2(1) = pin%sgp2
c0(1) = pin%c0; c1(1) = pin%c1; c2(1) = pin%c2; c3(1) = pin%c3; c4(1) = pin%c4
Qlj(1) = pin%lj /= 0; Qsrpmfi(1) = pin%srpmf /= 0; Qefpot(1) = .true.
call discretize(1, nnp, mnp, nxf1)
ep_out(1:3, 1:n) = ep(1:3, 1:n)
pout%sc1 = sc(1,1); pout%sc2 = sc(2,1); pout%sc3 = sc(3,1)
end subroutine

! ------------------------------------------------------------- map files
! C string -> Fortran character
subroutine c2f_path(cpath, fpath)
use iso_c_binding
implicit none
character(kind=c_char), intent(in) :: cpath(*)
character(len=*), intent(out) :: fpath
integer :: i
fpath = ' '
do i = 1, len(fpath)
  if (cpath(i) == c_null_char) exit
  fpath(i:i) = cpath(i)
enddo
end subroutine

! a grid file written by Fortran's own unformatted sequential write, in the record layout
! readphi.f90:43-44 reads
subroutine ref_write_map(cpath, nx, ny, nz, dcel, xc, yc, zc, rec2, vals) bind(C, name='ref_write_map')
use iso_c_binding
implicit none
character(kind=c_char), intent(in) :: cpath(*)
integer(c_int32_t), value :: nx, ny, nz
real(c_double), value :: dcel, xc, yc, zc
real(c_double), intent(in) :: rec2(6)
real(c_float), intent(in) :: vals(*)
character(len=1024) :: path
integer(4) :: i4x, i4y, i4z
real(8) :: d8, x8, y8, z8
integer :: u, i, n
call c2f_path(cpath, path)
i4x = nx; i4y = ny; i4z = nz; d8 = dcel; x8 = xc; y8 = yc; z8 = zc
n = nx*ny*nz
open(newunit=u, file=trim(path), form='unformatted', access='sequential', status='replace')
write(u) i4x, i4y, i4z, d8, x8, y8, z8
write(u) (real(rec2(i), 8), i = 1, 6)
write(u) (vals(i), i = 1, n)
close(u)
end subroutine

! READPHI (readphi.f90) on a file: kind 0 PHIX, 1 PHIV; the system box from box(6)
subroutine ref_readphi(cpath, kind, adjust, box, dims, geom, out) bind(C, name='ref_readphi')
use ref_harness
use grandmod
use gsbpmod
implicit none
character(kind=c_char), intent(in) :: cpath(*)
integer(c_int32_t), value :: kind, adjust
real(c_double), intent(in) :: box(6)
integer(c_int32_t), intent(out) :: dims(3)
real(c_double), intent(out) :: geom(4)
real(c_float), intent(out) :: out(*)
character(len=1024) :: path
integer :: u, nul, n
logical(1) :: adj
call c2f_path(cpath, path)
lx2m = box(1); lx2p = box(2); ly2m = box(3); ly2p = box(4); lz2m = box(5); lz2p = box(6)
Qnmcden = .false.
adj = adjust /= 0
open(newunit=u, file=trim(path), form='unformatted', access='sequential', status='old')
open(newunit=nul, file='/dev/null', status='old')
if (kind == 0) then
  call readphi(u, nul, 'PHIX', adj)
  dims = (/ nclx1, ncly1, nclz1 /)
  geom = (/ dcel1, xbcen1, ybcen1, zbcen1 /)
  n = nclx1*ncly1*nclz1
  out(1:n) = phix(1:n)
else
  call readphi(u, nul, 'PHIV', adj)
  dims = (/ nclx2, ncly2, nclz2 /)
  geom = (/ dcel2, xbcen2, ybcen2, zbcen2 /)
  n = nclx2*ncly2*nclz2
  out(1:n) = phiv(1:n)
endif
close(u)
close(nul)
end subroutine

! REPWALLS (repwalls.f90) on the PHIV map the last ref_readphi left (or the given one)
subroutine ref_repwalls(nx, ny, nz, map, walls) bind(C, name='ref_repwalls')
use ref_harness
use gsbpmod
implicit none
integer(c_int32_t), value :: nx, ny, nz, walls
integer(c_int8_t), intent(inout) :: map(*)
integer(1) :: w
integer :: n
nclx2 = nx; ncly2 = ny; nclz2 = nz
n = nx*ny*nz
Qnmcden = .false.
if (allocated(phiv)) deallocate(phiv)
allocate(phiv(n))
phiv(1:n) = map(1:n)
w = int(walls, 1)
call repwalls(w)
map(1:n) = phiv(1:n)
end subroutine

! READRFPAR (rfpar.f90:373-582) on one or two files; nion is set to the file count
subroutine ref_readrfpar(nfiles, cpath1, cpath2, single, adjust, box, dims, geom, radion_out, &
                         gs, gr) bind(C, name='ref_readrfpar')
use ref_harness
use grandmod
use gsbpmod
use nucleotmod
implicit none
integer(c_int32_t), value :: nfiles, single, adjust
character(kind=c_char), intent(in) :: cpath1(*), cpath2(*)
real(c_double), intent(in) :: box(6)
integer(c_int32_t), intent(out) :: dims(3)
real(c_double), intent(out) :: geom(4), radion_out(*)
real(c_float), intent(out) :: gs(*), gr(*)
character(len=1024) :: path1, path2
integer :: u(2), nul, n, k
logical(1) :: adj
call c2f_path(cpath1, path1)
call c2f_path(cpath2, path2)
lx2m = box(1); lx2p = box(2); ly2m = box(3); ly2p = box(4); lz2m = box(5); lz2p = box(6)
Qrfpsin = single /= 0
nion = nfiles
ndna = 0
atnam2(1:2) = (/ 'K   ', 'CL  ' /)
open(newunit=u(1), file=trim(path1), form='unformatted', access='sequential', status='old')
if (nfiles > 1) open(newunit=u(2), file=trim(path2), form='unformatted', access='sequential', status='old')
open(newunit=nul, file='/dev/null', status='old')
adj = adjust /= 0
if (allocated(gsrfen)) deallocate(gsrfen, greff)
call readrfpar(u(1:nfiles), int(nfiles), nul, adj)
dims = (/ nclx3, ncly3, nclz3 /)
geom = (/ dcel3, xbcen3, ybcen3, zbcen3 /)
n = nclx3*ncly3*nclz3*nfiles
gs(1:n) = gsrfen(1:n)
gr(1:n) = greff(1:n)
do k = 1, nfiles
  close(u(k))
enddo
close(nul)
radion_out(1) = 0.0
end subroutine
