! Test harness for the reference assembly path (test infrastructure only).
!
! Compiled together with the reference's own Fortran sources by
! oracle/Makefile into oracle/_ref/libwtref32.so (default REAL/COMPLEX) and
! oracle/_ref/libwtref64.so (-fdefault-real-8 promotion).  It sets the module
! globals the way the reference driver does and calls the reference routines
! unchanged:
!   readini.f90:81-84      ipml = INT(pmld/dx)+1, pmlf = 3 ln(1/pmlr)/(2 pmld^3)
!   modelio.f90:78-92      mu = vs^2 rho, lambda = vp^2 rho - 2 mu (complex)
!   elwave.f90:92-124      SREL3D_GETSIZE / SREL3D / SPARSE_ADJ2CRS
!   elwave.f90:151,164     omega = 2 pi f (+ i/tau)
!   elwave.f90:171         ASMBLE
!   haskgrn.f:31           HASKGRN (plane-wave depth profiles)
!   hetfldsrc.f90:17       HETFLDSRC (scattered-field source)
!   makeuh.f:1             MAKEUH (total field, receivers, hydrophone pressure)
!   dispersion.f:1         DISPERSION (attenuation: complex moduli)
! Only the oracle tests load the resulting library.
      SUBROUTINE wtref_graph_size(nx, ny, nz, numvert, nzero) &
                 BIND(C, NAME='wtref_graph_size')
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nx, ny, nz
      INTEGER(C_INT), INTENT(OUT) :: numvert, nzero
      INTEGER nedge
      CALL SREL3D_GETSIZE(nx, ny, nz, numvert, nedge)
      nzero = 2*nedge + numvert
      END SUBROUTINE

      SUBROUTINE wtref_graph(nx, ny, nz, irptr, jcptr) &
                 BIND(C, NAME='wtref_graph')
      USE ISO_C_BINDING
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nx, ny, nz
      INTEGER(C_INT), INTENT(OUT) :: irptr(*), jcptr(*)
      INTEGER, ALLOCATABLE :: xadj(:), adjncy(:)
      INTEGER numvert, nedge, nzero
      CALL SREL3D_GETSIZE(nx, ny, nz, numvert, nedge)
      ALLOCATE(xadj(numvert+1))
      ALLOCATE(adjncy(2*nedge))
      CALL SREL3D(nx, ny, nz, numvert, nedge, xadj, adjncy)
      nzero = 2*nedge + numvert
      CALL SPARSE_ADJ2CRS(nzero, numvert, xadj, adjncy, irptr, jcptr)
      DEALLOCATE(xadj)
      DEALLOCATE(adjncy)
      END SUBROUTINE

! Module globals of one model at one frequency, as readini / modelio /
! elwave set them; the background model (dap, mup, rhop) is a second copy
! of the given speeds.
      SUBROUTINE wtref_model(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, &
                             pmldi, pmlri, vp, vs, rh, vpb, vsb, rhb)
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, omega, dx, dy, dz, &
                               nx, ny, nz, freesurf
      USE PML_MODULE, ONLY : pmld, pmlr, pmlf, ipml
      USE INIT_MODULE, ONLY : usemin, tau
      IMPLICIT NONE
      INTEGER, INTENT(IN) :: nxi, nyi, nzi, fst, usem
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      REAL, INTENT(IN) :: vpb(nzi,nxi,nyi), vsb(nzi,nxi,nyi), rhb(nzi,nxi,nyi)
      COMPLEX vc
      REAL, PARAMETER :: pi = 3.141592653589793
      INTEGER ix, iy, iz
      nx = nxi
      ny = nyi
      nz = nzi
      dx = dxi
      dy = dxi
      dz = dxi
      freesurf(:) = .FALSE.
      freesurf(1) = (fst /= 0)
      usemin = (usem /= 0)
      tau = taui
      pmlr = pmlri
      pmld = pmldi
      ipml = INT(pmld/dx) + 1
      pmlf = 3.0*LOG(1.0/pmlr)/(2.0*pmld*pmld*pmld)
      IF (ALLOCATED(da)) DEALLOCATE(da, mu, rho, dap, mup, rhop)
      ALLOCATE(da(nz,nx,ny), mu(nz,nx,ny), rho(nz,nx,ny))
      ALLOCATE(dap(nz,nx,ny), mup(nz,nx,ny), rhop(nz,nx,ny))
      DO iy=1,ny
         DO ix=1,nx
            DO iz=1,nz
               rho(iz,ix,iy) = rh(iz,ix,iy)
               vc = CMPLX(vs(iz,ix,iy), 0.0)
               mu(iz,ix,iy) = vc*vc*rho(iz,ix,iy)
               vc = CMPLX(vp(iz,ix,iy), 0.0)
               da(iz,ix,iy) = vc*vc*rho(iz,ix,iy) - 2.0*mu(iz,ix,iy)
               rhop(iz,ix,iy) = rhb(iz,ix,iy)
               vc = CMPLX(vsb(iz,ix,iy), 0.0)
               mup(iz,ix,iy) = vc*vc*rhop(iz,ix,iy)
               vc = CMPLX(vpb(iz,ix,iy), 0.0)
               dap(iz,ix,iy) = vc*vc*rhop(iz,ix,iy) - 2.0*mup(iz,ix,iy)
            ENDDO
         ENDDO
      ENDDO
      omega = CMPLX(2.0*pi*freqi, 0.0)
      IF (tau /= 999.999) omega = omega + CMPLX(0.0, 1.0/tau)
      END SUBROUTINE

      SUBROUTINE wtref_assemble(nxi, nyi, nzi, dxi, freqi, taui, &
                                fst, usem, pmldi, pmlri, &
                                vp, vs, rh, nzero, irptr, jcptr, a, ierr) &
                 BIND(C, NAME='wtref_assemble')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, nx, ny, nz
      USE INTERFACE_MODULE, ONLY : asmble
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi, fst, usem, nzero
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      INTEGER(C_INT), INTENT(IN) :: irptr(*), jcptr(*)
      COMPLEX, INTENT(OUT) :: a(*)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      CALL wtref_model(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, pmldi, pmlri, &
                       vp, vs, rh, vp, vs, rh)
      CALL ASMBLE(6, nzero, nx, ny, nz, irptr, jcptr, a, ierr)
      DEALLOCATE(da, mu, rho, dap, mup, rhop)
      END SUBROUTINE

! HETFLDSRC (hetfldsrc.f90:17) for one model against a background model and
! background fields ub, vb, wb given on the grid ((nz,nx,ny), z fastest).
! The output buffer is sized for the reference's own node numbering
! (hetfldsrc.f90:1020, (iy-1)*nx*ny + (ix-1)*nz + iz), which equals the
! unknown order when ny = nz; the first 3*nx*ny*nz entries are returned.
      SUBROUTINE wtref_hetfldsrc(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, &
                                 pmldi, pmlri, vp, vs, rh, vpb, vsb, rhb, &
                                 ub, vb, wb, srcr, srci, sfld, ierr) &
                 BIND(C, NAME='wtref_hetfldsrc')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, nx, ny, nz
      USE WAVEFIELD_MODULE, ONLY : utbkr, vtbkr, wtbkr
      USE INTERFACE_MODULE, ONLY : hetfldsrc
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi, fst, usem
      REAL, INTENT(IN) :: dxi, freqi, taui, pmldi, pmlri, srcr, srci
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      REAL, INTENT(IN) :: vpb(nzi,nxi,nyi), vsb(nzi,nxi,nyi), rhb(nzi,nxi,nyi)
      COMPLEX, INTENT(IN) :: ub(nzi,nxi,nyi), vb(nzi,nxi,nyi), wb(nzi,nxi,nyi)
      COMPLEX, INTENT(OUT) :: sfld(3*nzi*nxi*nyi)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      COMPLEX, ALLOCATABLE :: sf(:)
      INTEGER ierr1
      CALL wtref_model(nxi, nyi, nzi, dxi, freqi, taui, fst, usem, pmldi, pmlri, &
                       vp, vs, rh, vpb, vsb, rhb)
      IF (ALLOCATED(utbkr)) DEALLOCATE(utbkr, vtbkr, wtbkr)
      ALLOCATE(utbkr(nz,nx,ny), vtbkr(nz,nx,ny), wtbkr(nz,nx,ny))
      utbkr = ub
      vtbkr = vb
      wtbkr = wb
      ALLOCATE(sf(3*((ny-1)*nx*ny + nx*nz)))
      CALL HETFLDSRC(CMPLX(srcr, srci), sf, 0, 0, ierr1)
      ierr = ierr1
      sfld = sf(1:3*nx*ny*nz)
      DEALLOCATE(sf, utbkr, vtbkr, wtbkr)
      DEALLOCATE(da, mu, rho, dap, mup, rhop)
      END SUBROUTINE

! HASKGRN (haskgrn.f:31) for source 1: the layer model goes into
! HASKELL_MODULE (top depths h1d, continuous and discrete speeds), the
! incidence angle and type into SRCPRM_MODULE ('pp' for P, 'ps' for S),
! the depth grid into MODEL_MODULE.
      SUBROUTINE wtref_haskgrn(nli, vp, vs, rh, top, vpd, vsd, ain1, iss, &
                               dfi, nzi, dzi, zorigi, ug, wg, ierr) &
                 BIND(C, NAME='wtref_haskgrn')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : dz, nz, zorig
      USE SRCPRM_MODULE, ONLY : ain, srctyp
      USE HASKELL_MODULE, ONLY : nl, ugrn1f, wgrn1f, vp1d, vs1d, rh1d, h1d, &
                                 vs1disc, vp1disc
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nli, iss, nzi
      REAL, INTENT(IN) :: vp(nli), vs(nli), rh(nli), top(nli), vpd(nli), vsd(nli)
      REAL, INTENT(IN) :: ain1, dfi, dzi, zorigi
      COMPLEX, INTENT(OUT) :: ug(nzi), wg(nzi)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      INTEGER ierr1
      nl = nli
      nz = nzi
      dz = dzi
      zorig = zorigi
      IF (ALLOCATED(vp1d)) DEALLOCATE(vp1d, vs1d, rh1d, h1d, vp1disc, vs1disc)
      ALLOCATE(vp1d(nl), vs1d(nl), rh1d(nl), h1d(nl), vp1disc(nl), vs1disc(nl))
      vp1d = vp
      vs1d = vs
      rh1d = rh
      h1d = top
      vp1disc = vpd
      vs1disc = vsd
      IF (ALLOCATED(ain)) DEALLOCATE(ain, srctyp)
      ALLOCATE(ain(1), srctyp(1))
      ain(1) = ain1
      srctyp(1) = 'pp'
      IF (iss /= 0) srctyp(1) = 'ps'
      IF (ALLOCATED(ugrn1f)) DEALLOCATE(ugrn1f, wgrn1f)
      ALLOCATE(ugrn1f(nz), wgrn1f(nz))
      ugrn1f = CMPLX(0.0, 0.0)
      wgrn1f = CMPLX(0.0, 0.0)
      CALL HASKGRN(dfi, 1, ierr1)
      ierr = ierr1
      ug = ugrn1f
      wg = wgrn1f
      DEALLOCATE(vp1d, vs1d, rh1d, h1d, vp1disc, vs1disc, ain, srctyp, ugrn1f, wgrn1f)
      END SUBROUTINE

! MAKEUH (makeuh.f:1-172) for one source at one frequency: the solution
! sfld (scattered field, unknown order), the background fields (nz,nx,ny),
! ng receivers (displacements, optional response factors rfac when recin
! is set) and nr hydrophones (pressure, with the model's lambda and mu from
! vp, vs, rh at the hydrophone nodes); every source-receiver pair active.
      SUBROUTINE wtref_makeuh(nxi, nyi, nzi, dxi, freqi, fst, vp, vs, rh, totf, &
                              sfld, ub, vb, wb, ngi, xgi, ygi, zgi, gsp, igr, &
                              reci, rfac, nri, xri, yri, zri, rsp, irr, &
                              ue, ve, we, pe, ierr) &
                 BIND(C, NAME='wtref_makeuh')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, dap, mup, rhop, nx, ny, nz, iom
      USE INIT_MODULE, ONLY : totfld
      USE SRCPRM_MODULE, ONLY : srctyp
      USE WAVEFIELD_MODULE, ONLY : utbkr, vtbkr, wtbkr, ufac, vfac, wfac
      USE RECPRM_MODULE, ONLY : xg, yg, zg, xr, yr, zr, sg, sr, receiver, &
                                uest, vest, west, utest, igreg, irreg, ng, nr, &
                                recin, gspread, rspread
      USE INTERFACE_MODULE, ONLY : makeuh
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi, fst, totf, ngi, igr, reci, nri, irr
      REAL, INTENT(IN) :: dxi, freqi, gsp, rsp
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      COMPLEX, INTENT(IN) :: sfld(3*nzi*nxi*nyi)
      COMPLEX, INTENT(IN) :: ub(nzi,nxi,nyi), vb(nzi,nxi,nyi), wb(nzi,nxi,nyi)
      REAL, INTENT(IN) :: xgi(*), ygi(*), zgi(*), xri(*), yri(*), zri(*)
      COMPLEX, INTENT(IN) :: rfac(*)
      COMPLEX, INTENT(OUT) :: ue(*), ve(*), we(*), pe(*)
      INTEGER(C_INT), INTENT(OUT) :: ierr
      INTEGER ierr1
      COMPLEX, ALLOCATABLE :: sf(:)
      CALL wtref_model(nxi, nyi, nzi, dxi, freqi, 999.999, fst, 1, 10.0*dxi, 1.0e-3, &
                       vp, vs, rh, vp, vs, rh)
      iom = 1
      totfld = (totf /= 0)
      IF (ALLOCATED(srctyp)) DEALLOCATE(srctyp)
      ALLOCATE(srctyp(1))
      srctyp(1) = 'pp'
      IF (ALLOCATED(utbkr)) DEALLOCATE(utbkr, vtbkr, wtbkr)
      ALLOCATE(utbkr(nz,nx,ny), vtbkr(nz,nx,ny), wtbkr(nz,nx,ny))
      utbkr = ub
      vtbkr = vb
      wtbkr = wb
      IF (ALLOCATED(ufac)) DEALLOCATE(ufac, vfac, wfac)
      ALLOCATE(ufac(nz,nx,ny), vfac(nz,nx,ny), wfac(nz,nx,ny))
      ng = ngi
      nr = nri
      igreg = igr
      irreg = irr
      gspread = gsp
      rspread = rsp
      recin = reci
      IF (ALLOCATED(xg)) DEALLOCATE(xg, yg, zg, sg, receiver, uest, vest, west)
      ALLOCATE(xg(MAX(ng,1)), yg(MAX(ng,1)), zg(MAX(ng,1)), sg(1,MAX(ng,1)))
      ALLOCATE(receiver(MAX(ng,1),1), uest(1,MAX(ng,1),1), vest(1,MAX(ng,1),1), west(1,MAX(ng,1),1))
      IF (ng > 0) THEN
         xg(1:ng) = xgi(1:ng)
         yg(1:ng) = ygi(1:ng)
         zg(1:ng) = zgi(1:ng)
         receiver(1:ng,1) = rfac(1:ng)
      ENDIF
      sg = 1
      IF (ALLOCATED(xr)) DEALLOCATE(xr, yr, zr, sr, utest)
      ALLOCATE(xr(MAX(nr,1)), yr(MAX(nr,1)), zr(MAX(nr,1)), sr(1,MAX(nr,1)), utest(1,MAX(nr,1),1))
      IF (nr > 0) THEN
         xr(1:nr) = xri(1:nr)
         yr(1:nr) = yri(1:nr)
         zr(1:nr) = zri(1:nr)
      ENDIF
      sr = 1
      ALLOCATE(sf(3*nz*nx*ny))
      sf = sfld
      CALL MAKEUH(1, sf, ierr1)
      DEALLOCATE(sf)
      ierr = ierr1
      IF (ng > 0) THEN
         ue(1:ng) = uest(1,1:ng,1)
         ve(1:ng) = vest(1,1:ng,1)
         we(1:ng) = west(1,1:ng,1)
      ENDIF
      IF (nr > 0) pe(1:nr) = utest(1,1:nr,1)
      DEALLOCATE(utbkr, vtbkr, wtbkr, ufac, vfac, wfac, srctyp)
      DEALLOCATE(xg, yg, zg, sg, receiver, uest, vest, west, xr, yr, zr, sr, utest)
      DEALLOCATE(da, mu, rho, dap, mup, rhop)
      END SUBROUTINE

! DISPERSION (dispersion.f:1-131) on a grid of speeds vp, vs at the base
! frequency, inverse quality factors qpi, qsi (the reference's qp, qs) and
! densities; da and mu start from the lossless moduli as modelio.f90:78-92
! leaves them.
      SUBROUTINE wtref_dispersion(nxi, nyi, nzi, vp, vs, qpi, qsi, rh, fbase, omr, lam, mu1) &
                 BIND(C, NAME='wtref_dispersion')
      USE ISO_C_BINDING
      USE MODEL_MODULE, ONLY : da, mu, rho, vpr, vsr, qp, qs, omega, freqbase, nx, ny, nz
      USE INIT_MODULE, ONLY : ncom
      IMPLICIT NONE
      INTEGER(C_INT), INTENT(IN) :: nxi, nyi, nzi
      REAL, INTENT(IN) :: vp(nzi,nxi,nyi), vs(nzi,nxi,nyi), qpi(nzi,nxi,nyi), qsi(nzi,nxi,nyi), rh(nzi,nxi,nyi)
      REAL, INTENT(IN) :: fbase, omr
      COMPLEX, INTENT(OUT) :: lam(nzi,nxi,nyi), mu1(nzi,nxi,nyi)
      COMPLEX vc
      INTEGER ix, iy, iz
      nx = nxi
      ny = nyi
      nz = nzi
      ncom = 0
      IF (ALLOCATED(da)) DEALLOCATE(da, mu, rho)
      IF (ALLOCATED(vpr)) DEALLOCATE(vpr, vsr, qp, qs)
      ALLOCATE(da(nz,nx,ny), mu(nz,nx,ny), rho(nz,nx,ny), vpr(nz,nx,ny), vsr(nz,nx,ny), qp(nz,nx,ny), qs(nz,nx,ny))
      vpr = vp
      vsr = vs
      qp = qpi
      qs = qsi
      rho = rh
      DO iy=1,ny
         DO ix=1,nx
            DO iz=1,nz
               vc = CMPLX(vs(iz,ix,iy), 0.0)
               mu(iz,ix,iy) = vc*vc*rho(iz,ix,iy)
               vc = CMPLX(vp(iz,ix,iy), 0.0)
               da(iz,ix,iy) = vc*vc*rho(iz,ix,iy) - 2.0*mu(iz,ix,iy)
            ENDDO
         ENDDO
      ENDDO
      omega = CMPLX(omr, 0.0)
      freqbase = fbase
      CALL DISPERSION()
      lam = da
      mu1 = mu
      DEALLOCATE(da, mu, rho, vpr, vsr, qp, qs)
      END SUBROUTINE
