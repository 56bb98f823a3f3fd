!-----------------------------------------------------------------------------
! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! Drives the reference's own path routines (compiled in place from
! /root/reference/src by oracle/Makefile) for one level, so that the C oracle
! (oracle/avism_oracle.c) can be pinned against them.  This program is ours; it
! restates the single-level flow of the reference driver:
!   particle mode : avism.f90:657-700 (tree), 1177-1257 (fields),
!                   1371-1414 (mean density)
!   grid mode     : reader.f90:523-552 and avism.f90:1299-1342 (FLAG_DATA=2)
!   both          : avism.f90:1633-1973 (marking, VOIDFIND, merge, post)
!
! Usage:  ref_harness_pX.x <jobdir>
!   <jobdir>/job.in      list-directed parameters (see READ statements below)
!   <jobdir>/parts.bin   (particle mode) x,y,z,vx,vy,vz,m  float32 [np] each
!   <jobdir>/fields.bin  (grid mode) delta,vx,vy,vz  float32 [n^3] each
!   <jobdir>/keys.bin    (mode 3) float32 keys for INDEXX
!   <jobdir>/marcap.bin  (optional) MARCAP int32 [n^3] and <jobdir>/reqp.bin
!                        REQP float32 [nvoid_max] of a parent level, as the
!                        reference driver hands them to the next level
!                        (avism.f90:2123-2161); without them MARCAP = -1 and
!                        REQP = 0 (avism.f90:957-964)
!   <jobdir>/smask.bin   (optional) SMASK int8 [n^3] on the inner grid; the
!                        periodic shell stays 0 as in survey mode
!                        (avism.f90:866-904); without it SMASK = 1
! Outputs <jobdir>/out_*.bin as raw stream files.
!-----------------------------------------------------------------------------
PROGRAM REF_HARNESS
   USE COMMONDATA
   USE COSMOKDTREE
   USE PARTICLES
   USE VOIDFINDING
   IMPLICIT NONE
   CHARACTER(LEN=512) :: JOB
   INTEGER :: MODE, N, NP4, KNEIGH, NPLUS, LOW1, LOW2, IX, JY, KZ, I, IV, IND, NVOID, NVOID2
   INTEGER :: NVOIDT
   LOGICAL :: HAVE_MARCAP, HAVE_MASK, HAVE_RANDS
   INTEGER*4 :: NRAND, KNEIGH2
   REAL*4, ALLOCATABLE :: RANDX(:), RANDY(:), RANDZ(:), MASSRAND(:), HRAND(:), U1RAND(:,:,:), RPTS(:,:)
   TYPE(KDTreeNode), POINTER :: RTREE
   REAL*8 :: DSUM, RSUM, ALPHA, BASS8
   REAL*4 :: RATIOFLOOR, RATIOMAX, MEANRATIO
   INTEGER(KIND=8) :: I8, CNT8
   REAL*4 :: ACHE, OMEGA0, ZETA, ACHE0, RODO, RE0, ROTE, RETE, HTE, LADO, SIZING, MEANDENS
   REAL*4 :: DXX, RX1
   REAL*8 :: SUM8, VOLM, VOLT_CLEAN
   REAL*4, ALLOCATABLE :: UBAS(:,:,:), U2C(:,:,:), U3C(:,:,:), U4C(:,:,:), U1C(:,:,:)
   INTEGER*1, ALLOCATABLE :: SMASK(:,:,:)
   REAL, ALLOCATABLE :: HPART(:), PART_DENS(:), TREEPTS(:,:)
   REAL :: LPER(3)
   TYPE(KDTreeNode), POINTER :: TREE
   INTEGER, ALLOCATABLE :: INDICE(:), INDICE2(:), UVOID(:), NCELLV(:)
   REAL*4, ALLOCATABLE :: VOLNEW(:), GXC(:), GYC(:), GZC(:), XC(:), YC(:), ZC(:)
   REAL*4, ALLOCATABLE :: EPS(:), IP(:), UMEAN(:), MTOT(:), REQ(:)
   REAL*4, ALLOCATABLE :: REQP(:)

   CALL GET_COMMAND_ARGUMENT(1, JOB)
   OPEN(1, FILE=TRIM(JOB)//'/job.in', STATUS='OLD', ACTION='READ')
   READ(1,*) MODE            ! 1 = particles, 2 = grid
   READ(1,*) N, LADO0, FLAG_PERIOD
   READ(1,*) DENS_THRE, DENS_THRE2, GRAD_THRE, DIV_THRE, RMIN, RMIN_SUB, NVOID_MAX
   READ(1,*) ACHE, OMEGA0, ZETA
   READ(1,*) KNEIGH, FLAG_VEL_INTERP, FLAG_DENS_INTERP, NP4
   CLOSE(1)

   IF (MODE .EQ. 3) THEN
      ! INDEXX on a key array (num.f90:2-84)
      CALL RUN_INDEXX(NP4)
      STOP
   ENDIF
   IF (MODE .EQ. 5) THEN
      ! READ_MASCLET (reader.f90:86-405) on simu_masclet/ of the working directory
      CALL RUN_READ_MASCLET(N, NP4)
      STOP
   ENDIF
   IF (MODE .EQ. 4) THEN
      ! MASCLET AMR gas onto a level's uniform grid (SMOOTH_2, DIVER_FINA_GAS, VMESH)
      CALL RUN_AMR(N)
      STOP
   ENDIF

   PI = DACOS(-1.D0)
   ACHE0 = ACHE*3.66D-3
   RODO = OMEGA0*3.D0*ACHE0**2
   RE0 = 1.0/10.98
   ROTE = RODO*(1.0+ZETA)**3
   RETE = RE0/(1.0+ZETA)
   HTE = OMEGA0*(((RE0/RETE)**3) - 1.0) + 1.0    ! avism.f90:601-602
   HTE = ACHE*SQRT(HTE)

   ! ---- grid set-up (avism.f90:780-865)
   IF (FLAG_PERIOD .EQ. 0) THEN
      NPLUS = 0; LOW1 = 1; LOW2 = N; LADO0PLUS = LADO0
      ALLOCATE(RADX(0:N+1), RADY(0:N+1), RADZ(0:N+1))
   ELSE
      SIZING = 50.
      DX = LADO0/N
      SIZING = MIN(SIZING, LADO0/2.)
      NPLUS = MAX(N/8, INT(SIZING/DX))
      LOW1 = -NPLUS+1; LOW2 = N+NPLUS
      LADO0PLUS = LADO0 + 2.*NPLUS*DX
      ALLOCATE(RADX(LOW1:LOW2), RADY(LOW1:LOW2), RADZ(LOW1:LOW2))
   ENDIF
   LADO = LADO0-(LADO0/N)
   CALL MALLA(N,N,N,LADO)
   DXX = DX
   RX1 = RADX(1)
   IF (FLAG_PERIOD .EQ. 1) THEN
      DO IX = LOW1, LOW2
         IF (IX .LT. 1) THEN
            RADX(IX) = RADX(1) - (1-IX)*DXX
         ELSE IF (IX .GT. N) THEN
            RADX(IX) = RADX(N) + (IX-N)*DXX
         ENDIF
      ENDDO
      RADY = RADX
      RADZ = RADX
   ENDIF
   ALLOCATE(SMASK(LOW1:LOW2,LOW1:LOW2,LOW1:LOW2))
   SMASK = 1
   INQUIRE(FILE=TRIM(JOB)//'/smask.bin', EXIST=HAVE_MASK)
   IF (HAVE_MASK) THEN
      SMASK = 0
      OPEN(2, FILE=TRIM(JOB)//'/smask.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) SMASK(1:N,1:N,1:N)
      CLOSE(2)
   ENDIF
   ALLOCATE(UBAS(N,N,N), U1C(N,N,N), U2C(N,N,N), U3C(N,N,N), U4C(N,N,N))
   U1C = 0.; U2C = 0.; U3C = 0.; U4C = 0.

   ALLOCATE(INICIOX(NVOID_MAX), INICIOY(NVOID_MAX), INICIOZ(NVOID_MAX))
   ALLOCATE(FINALX(NVOID_MAX), FINALY(NVOID_MAX), FINALZ(NVOID_MAX))
   ALLOCATE(ICX(NVOID_MAX), ICY(NVOID_MAX), ICZ(NVOID_MAX), VOL(NVOID_MAX), FATHER(NVOID_MAX))
   ALLOCATE(RINIXCO(NVOID_MAX), RINIYCO(NVOID_MAX), RINIZCO(NVOID_MAX))
   ALLOCATE(RFINXCO(NVOID_MAX), RFINYCO(NVOID_MAX), RFINZCO(NVOID_MAX))
   ALLOCATE(MARCAP(N,N,N), REQP(NVOID_MAX))
   MARCAP = -1
   REQP = 0.
   FATHER = 0
   INQUIRE(FILE=TRIM(JOB)//'/marcap.bin', EXIST=HAVE_MARCAP)
   IF (HAVE_MARCAP) THEN
      OPEN(2, FILE=TRIM(JOB)//'/marcap.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) MARCAP
      CLOSE(2)
      OPEN(2, FILE=TRIM(JOB)//'/reqp.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) REQP
      CLOSE(2)
   ENDIF

   ALLOCATE(U1CO(LOW1:LOW2,LOW1:LOW2,LOW1:LOW2), DIVERCO(LOW1:LOW2,LOW1:LOW2,LOW1:LOW2))

   IF (MODE .EQ. 1) THEN
      ! ---- particles (reader.f90:410-465 then avism.f90:657-1257)
      NPARTT = NP4
      ALLOCATE(RXPA(NPARTT), RYPA(NPARTT), RZPA(NPARTT), U2PA(NPARTT), U3PA(NPARTT), U4PA(NPARTT), MASAP(NPARTT))
      OPEN(2, FILE=TRIM(JOB)//'/parts.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) RXPA, RYPA, RZPA, U2PA, U3PA, U4PA, MASAP
      CLOSE(2)
      MASAP = MASAP/UM
      U2PA = U2PA/UV
      U3PA = U3PA/UV
      U4PA = U4PA/UV
      ALLOCATE(HPART(NPARTT), TREEPTS(NPARTT,3), PART_DENS(NPARTT))
      HPART = 0.
      PART_DENS = 0.
      TREEPTS(:,1) = RXPA; TREEPTS(:,2) = RYPA; TREEPTS(:,3) = RZPA
#if periodic == 1
      LPER = [LADO0, LADO0, LADO0]
      TREE => build_kdtree(TREEPTS, LPER)
#else
      TREE => build_kdtree(TREEPTS)
#endif
      IF (FLAG_VEL_INTERP .EQ. 0) CALL VEL_INTERP_TSC(N,N,N,NPARTT,RXPA,RYPA,RZPA,U2PA,U3PA,U4PA,U2C,U3C,U4C)
      IF (FLAG_VEL_INTERP .EQ. 1) THEN
         CALL PPART_DENS(KNEIGH,NPARTT,TREE,MASAP,PART_DENS)
         CALL VVEL_INTERP_SPH_VW(KNEIGH,N,N,N,NPARTT,TREE,HPART,PART_DENS,MASAP,U2PA,U3PA,U4PA, &
                                 SMASK(1:N,1:N,1:N),U2C,U3C,U4C)
         CALL FILL_H_ZEROS(KNEIGH,NPARTT,RXPA,RYPA,RZPA,TREE,HPART)
      ENDIF
      IF (FLAG_DENS_INTERP .EQ. 0) CALL DENS_INTERP_TSC(N,N,N,NPARTT,RXPA,RYPA,RZPA,MASAP,U1C)
      IF (FLAG_DENS_INTERP .EQ. 1) THEN
         IF (FLAG_VEL_INTERP .LE. 0) THEN
            CALL H_DISTANCE(KNEIGH,N,N,N,NPARTT,TREE,SMASK(1:N,1:N,1:N),HPART)
            CALL FILL_H_ZEROS(KNEIGH,NPARTT,RXPA,RYPA,RZPA,TREE,HPART)
         ENDIF
         CALL DDENS_INTERP_SPH(N,N,N,NPARTT,RXPA,RYPA,RZPA,HPART,MASAP,SMASK(1:N,1:N,1:N),U1C)
      ENDIF
      OPEN(3, FILE=TRIM(JOB)//'/out_mass.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) U1C
      CLOSE(3)
      OPEN(3, FILE=TRIM(JOB)//'/out_parts.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) PART_DENS, HPART
      CLOSE(3)
      U1C = U1C / (DXX*DXX*DXX*RETE**3)
      U1C = U1C / ROTE
      CALL DIVER_UNIFORM(N,N,N,DXX,DXX,DXX,U2C,U3C,U4C,UBAS,FLAG_PERIOD)
      U1CO(1:N,1:N,1:N) = U1C
      U1CO = U1CO*ROTE
      INQUIRE(FILE=TRIM(JOB)//'/rands.bin', EXIST=HAVE_RANDS)
      IF (HAVE_RANDS) THEN
         ! survey mode (avism.f90:1420-1590); rands.bin: NRAND, then x, y, z, m
         OPEN(2, FILE=TRIM(JOB)//'/rands.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
         READ(2) NRAND
         ALLOCATE(RANDX(NRAND), RANDY(NRAND), RANDZ(NRAND), MASSRAND(NRAND), U1RAND(N,N,N))
         READ(2) RANDX, RANDY, RANDZ, MASSRAND
         CLOSE(2)
         MASSRAND = MASSRAND/UM
         U1RAND = 0.
         IF (FLAG_DENS_INTERP .EQ. 0) THEN
            CALL DENS_INTERP_TSC(N,N,N,INT(NRAND,KIND=8),RANDX,RANDY,RANDZ,MASSRAND,U1RAND)
         ELSE
            ALLOCATE(HRAND(NRAND), RPTS(NRAND,3))
            HRAND = 0.
            RPTS(:,1) = RANDX(:)
            RPTS(:,2) = RANDY(:)
            RPTS(:,3) = RANDZ(:)
#if periodic == 1
            RTREE => build_kdtree(RPTS, LPER)
#else
            RTREE => build_kdtree(RPTS)
#endif
            KNEIGH2 = MIN(8*KNEIGH, 256)
            CALL H_DISTANCE(KNEIGH2,N,N,N,INT(NRAND,KIND=8),RTREE,SMASK(1:N,1:N,1:N),HRAND)
            CALL FILL_H_ZEROS(KNEIGH2,INT(NRAND,KIND=8),RANDX,RANDY,RANDZ,RTREE,HRAND)
            CALL DDENS_INTERP_SPH(N,N,N,INT(NRAND,KIND=8),RANDX,RANDY,RANDZ,HRAND,MASSRAND, &
                                  SMASK(1:N,1:N,1:N),U1RAND)
         ENDIF
         DSUM = 0.0D0
         RSUM = 0.0D0
         DO KZ=1,N
            DO JY=1,N
               DO IX=1,N
                  IF (SMASK(IX,JY,KZ) .EQ. 0) CYCLE
                  DSUM = DSUM + REAL(U1CO(IX,JY,KZ), KIND=8)
                  RSUM = RSUM + REAL(U1RAND(IX,JY,KZ), KIND=8)
               ENDDO
            ENDDO
         ENDDO
         ALPHA = DSUM/RSUM
         RATIOFLOOR = 1E-4
         RATIOMAX = 1000.0D0
         DO KZ=1,N
            DO JY=1,N
               DO IX=1,N
                  IF (SMASK(IX,JY,KZ) .EQ. 0) CYCLE
                  IF (U1RAND(IX,JY,KZ) > 0.0D0) THEN
                     BASS8 = U1CO(IX,JY,KZ)/(ALPHA*U1RAND(IX,JY,KZ))
                  ELSE
                     BASS8 = RATIOFLOOR
                  ENDIF
                  BASS8 = MAX(REAL(RATIOFLOOR, KIND=8), BASS8)
                  BASS8 = MIN(REAL(RATIOMAX, KIND=8), BASS8)
                  U1CO(IX,JY,KZ) = BASS8
               ENDDO
            ENDDO
         ENDDO
         MEANRATIO = SUM(REAL(U1CO(1:N,1:N,1:N),8), MASK=SMASK(1:N,1:N,1:N)>0) / &
                     REAL(COUNT(SMASK(1:N,1:N,1:N)>0),8)
         WHERE (SMASK(1:N,1:N,1:N)>0) U1CO(1:N,1:N,1:N) = U1CO(1:N,1:N,1:N) / MEANRATIO
         MEANDENS = ROTE * (UM / UL**3)
         OPEN(3, FILE=TRIM(JOB)//'/out_rands.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
         WRITE(3) U1RAND
         CLOSE(3)
      ELSE
      SUM8 = 0.D0
      CNT8 = 0
      DO KZ=1,N
         DO JY=1,N
            DO IX=1,N
               SUM8 = SUM8 + REAL(U1CO(IX,JY,KZ), KIND=8)
               CNT8 = CNT8 + 1
            ENDDO
         ENDDO
      ENDDO
      MEANDENS = SUM8 / REAL(CNT8, KIND=8)
      U1CO = U1CO / MEANDENS
      MEANDENS = MEANDENS * (UM / UL**3)
      ENDIF
   ELSE
      ! ---- grid input (reader.f90:523-552, FLAG_DATA=2)
      OPEN(2, FILE=TRIM(JOB)//'/fields.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) U1C, U2C, U3C, U4C
      CLOSE(2)
      U1C = U1C + 1.0
      U2C = U2C/UV
      U3C = U3C/UV
      U4C = U4C/UV
      CALL DIVER_UNIFORM(N,N,N,DXX,DXX,DXX,U2C,U3C,U4C,UBAS,FLAG_PERIOD)
      U1CO(1:N,1:N,1:N) = U1C
      MEANDENS = ROTE * (UM / UL**3)
   ENDIF
   OPEN(3, FILE=TRIM(JOB)//'/out_vel.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
   WRITE(3) U2C, U3C, U4C
   CLOSE(3)

   DIVERCO(1:N,1:N,1:N) = UBAS
   ! FLAG_VEL_INTERP = -1 here: velocities not available, FLAG_DIV = -1
   ! (avism.f90:1259-1263, 1646-1650), on the inner grid
   IF (MODE .EQ. 1 .AND. FLAG_VEL_INTERP .EQ. -1) &
      DIVERCO(1:N,1:N,1:N) = - OMEGA0**(0.6)*RETE*HTE*(U1CO(1:N,1:N,1:N) - 1.0)
   IF (FLAG_PERIOD .EQ. 1) THEN
      CALL REAL_MAKE_PERIODIC(N,N,N,NPLUS,U1CO)
      CALL REAL_MAKE_PERIODIC(N,N,N,NPLUS,DIVERCO)
   ENDIF
   ALLOCATE(MARCAP2(LOW1:LOW2,LOW1:LOW2,LOW1:LOW2))
   MARCAP2(1:N,1:N,1:N) = MARCAP
   IF (FLAG_PERIOD .EQ. 1) CALL INT_MAKE_PERIODIC(N,N,N,NPLUS,MARCAP2)
   CALL MARK_ALL(LOW1,LOW2,SMASK)
   CALL VOIDFIND(0,LOW1,LOW2,DXX,DXX,DXX,RX1,RX1,RX1,SMASK,NVOID_MAX,REQP,NVOID)
   DEALLOCATE(FLAGV, FLAG_SUB)

   ALLOCATE(INDICE(NVOID), INDICE2(NVOID))
   CALL INDEXX(NVOID, VOL(1:NVOID), INDICE2)
   DO I = 1, NVOID
      INDICE(NVOID+1-I) = INDICE2(I)
   ENDDO
   ALLOCATE(GXC(NVOID),GYC(NVOID),GZC(NVOID),XC(NVOID),YC(NVOID),ZC(NVOID),UVOID(NVOID),VOLNEW(NVOID))
   ALLOCATE(UMEAN(NVOID),MTOT(NVOID),NCELLV(NVOID),EPS(NVOID),IP(NVOID),REQ(NVOID))
   ALLOCATE(MARCA(LOW1:LOW2,LOW1:LOW2,LOW1:LOW2))
   XC = 0.; YC = 0.; ZC = 0.; UMEAN = 0.; MTOT = 0.; REQ = 0.
   CALL MERGE_VOID(NVOID,INDICE,LOW1,LOW2,DXX,DXX,DXX,VOLNEW,UVOID,GXC,GYC,GZC,NVOID2)

   ! ---- post-processing (single level), restating avism.f90:1796-1973
   NCELLV = 0
   DO KZ=LOW1,LOW2
      DO JY=LOW1,LOW2
         DO IX=LOW1,LOW2
            IND = MARCA(IX,JY,KZ)
            IF (IND .GT. 0) THEN
               IF (UVOID(IND) .EQ. -1) NCELLV(IND) = NCELLV(IND) + 1
            ENDIF
         ENDDO
      ENDDO
   ENDDO
   DO I = 1, NVOID
      IND = INDICE(I)
      IF (UVOID(IND) .GE. 0) CYCLE
      IF (NCELLV(IND) .LT. NCELL_MIN) CALL DROP_VOID(IND)
   ENDDO
   CALL VOID_SHAPE(LOW1,LOW2,NVOID,INDICE,NCELLV,UVOID,GXC,GYC,GZC,EPS,IP)
   CALL SET_VOLNEW()
   CALL RMIN_FILTER()
   DO IV = 1, NVOID
      IND = INDICE(IV)
      IF (UVOID(IND) .GE. 0) CYCLE
      XC(IND) = RADX(1)+(ICX(IND)-1)*DXX
      YC(IND) = RADY(1)+(ICY(IND)-1)*DXX
      ZC(IND) = RADZ(1)+(ICZ(IND)-1)*DXX
   ENDDO
   CALL SIMPLE_CONNECTION_FAST(LOW1,LOW2,MARCA)
   IF (FLAG_PERIOD .EQ. 1) THEN
      CALL VOID_PERIODIC(NVOID,INDICE,UVOID,GXC,GYC,GZC,VOLNEW,N,N,N,DXX,DXX,DXX,LOW1,LOW2)
   ENDIF
   NCELLV = 0
   UMEAN = 0.
   MTOT = 0.
   VOLT_CLEAN = 0.
   DO KZ=1,N
      DO JY=1,N
         DO IX=1,N
            IND = MARCA(IX,JY,KZ)
            IF (IND .GT. 0) THEN
               IF (UVOID(IND) .EQ. -1) THEN
                  NCELLV(IND) = NCELLV(IND)+1
                  UMEAN(IND) = UMEAN(IND)+U1CO(IX,JY,KZ)
                  MTOT(IND) = MTOT(IND)+U1CO(IX,JY,KZ)*MEANDENS*DXX*DXX*DXX*RETE**3
                  VOLT_CLEAN = VOLT_CLEAN+DBLE(DXX*DXX*DXX)
               ENDIF
            ENDIF
         ENDDO
      ENDDO
   ENDDO
   WHERE (NCELLV .GT. 0)
      UMEAN = UMEAN / REAL(NCELLV)
   END WHERE
   CALL SET_VOLNEW()
   CALL RMIN_FILTER()
   NVOIDT = COUNT(UVOID .EQ. -1)
   DO IV = 1, NVOID
      IND = INDICE(IV)
      IF (UVOID(IND) .NE. -1) CYCLE
      VOLM = VOLNEW(IND)
      REQ(IND) = ((3.*VOLM)/(4.*PI))**(1./3.)
   ENDDO

   ! ---- outputs
   OPEN(3, FILE=TRIM(JOB)//'/out_head.txt', STATUS='REPLACE')
   WRITE(3,*) NVOID, NVOID2, NVOIDT, LOW1, LOW2, NPLUS
   WRITE(3,'(ES26.17E3)') VOLT_CLEAN
   WRITE(3,'(ES26.17E3)') MEANDENS
   CLOSE(3)
   OPEN(3, FILE=TRIM(JOB)//'/out_fields.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
   WRITE(3) U1CO(1:N,1:N,1:N), DIVERCO(1:N,1:N,1:N)
   CLOSE(3)
   OPEN(3, FILE=TRIM(JOB)//'/out_marca.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
   WRITE(3) MARCA
   CLOSE(3)
   OPEN(3, FILE=TRIM(JOB)//'/out_cubes.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
   IF (NVOID .GT. 0) THEN
      WRITE(3) INICIOX(1:NVOID), INICIOY(1:NVOID), INICIOZ(1:NVOID), FINALX(1:NVOID), FINALY(1:NVOID), &
               FINALZ(1:NVOID), ICX(1:NVOID), ICY(1:NVOID), ICZ(1:NVOID), FATHER(1:NVOID), INDICE, UVOID, NCELLV
      WRITE(3) RINIXCO(1:NVOID), RINIYCO(1:NVOID), RINIZCO(1:NVOID), RFINXCO(1:NVOID), RFINYCO(1:NVOID), &
               RFINZCO(1:NVOID), VOL(1:NVOID), GXC, GYC, GZC, XC, YC, ZC, VOLNEW, UMEAN, MTOT, EPS, IP, REQ
   ENDIF
   CLOSE(3)

CONTAINS

   SUBROUTINE RUN_INDEXX(NK)
      INTEGER :: NK
      REAL*4, ALLOCATABLE :: KEYS(:)
      INTEGER, ALLOCATABLE :: IDX(:)
      ALLOCATE(KEYS(NK), IDX(NK))
      OPEN(2, FILE=TRIM(JOB)//'/keys.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) KEYS
      CLOSE(2)
      CALL INDEXX(NK, KEYS, IDX)
      OPEN(3, FILE=TRIM(JOB)//'/out_indexx.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) IDX
      CLOSE(3)
   END SUBROUTINE RUN_INDEXX

   ! MASCLET gas of a base grid of NB^3 cells and its level-1 patches onto the
   ! uniform grid of level IR (amr.in: IR, NPATCH), as READ_MASCLET
   ! (reader.f90:86-405) leaves the arrays and avism.f90:366-382, 750, 1064-1146
   ! use them: IR < 0 SMOOTH_2, IR = 0 the base grid, then DIVER_UNIFORM;
   ! IR >= 1 DIVER_FINA_GAS and VMESH.
   !   amr_base.bin    delta, vx, vy, vz  float32 [NB^3] each
   !   amr_pint.bin    int32 [6, NPATCH]: PATCHNX, PATCHNY, PATCHNZ, PATCHX, PATCHY, PATCHZ
   !   amr_preal.bin   float32 [3, NPATCH]: PATCHRX, PATCHRY, PATCHRZ
   !   amr_patch.bin   per patch delta, vx, vy, vz  float32 [nx*ny*nz] each
   ! Output out_amr.bin: U1GCO, DIVERGCO  float32 [NXX^3] each.
   SUBROUTINE RUN_AMR(NB)
      INTEGER :: NB
      INTEGER :: IRL, NPA, NXL, IP, N1, N2, N3, D1, D2, D3, D4
      INTEGER, ALLOCATABLE :: PINT(:,:)
      REAL*4, ALLOCATABLE :: PREAL(:,:), DIVB(:,:,:)
      REAL*4 :: LADOB, DXL, RXL1
      OPEN(1, FILE=TRIM(JOB)//'/amr.in', STATUS='OLD', ACTION='READ')
      READ(1,*) IRL, NPA
      CLOSE(1)
      NHYX = NB; NHYY = NB; NHYZ = NB
      NL2 = 1; NLEVELS = 1; NPALEV = MAX(NPA, 1)
      ALLOCATE(NPATCH(0:NLEVELS), PARE(NPALEV))
      ALLOCATE(PATCHNX(NPALEV), PATCHNY(NPALEV), PATCHNZ(NPALEV))
      ALLOCATE(PATCHX(NPALEV), PATCHY(NPALEV), PATCHZ(NPALEV))
      ALLOCATE(PATCHRX(NPALEV), PATCHRY(NPALEV), PATCHRZ(NPALEV))
      NPATCH = 0; NPATCH(1) = NPA; PARE = 0
      PATCHNX = 0; PATCHNY = 0; PATCHNZ = 0; PATCHX = 0; PATCHY = 0; PATCHZ = 0
      PATCHRX = 0.; PATCHRY = 0.; PATCHRZ = 0.
      IF (NPA .GT. 0) THEN
         ALLOCATE(PINT(6, NPA), PREAL(3, NPA))
         OPEN(2, FILE=TRIM(JOB)//'/amr_pint.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
         READ(2) PINT
         CLOSE(2)
         OPEN(2, FILE=TRIM(JOB)//'/amr_preal.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
         READ(2) PREAL
         CLOSE(2)
         PATCHNX(1:NPA) = PINT(1,:); PATCHNY(1:NPA) = PINT(2,:); PATCHNZ(1:NPA) = PINT(3,:)
         PATCHX(1:NPA) = PINT(4,:); PATCHY(1:NPA) = PINT(5,:); PATCHZ(1:NPA) = PINT(6,:)
         PATCHRX(1:NPA) = PREAL(1,:); PATCHRY(1:NPA) = PREAL(2,:); PATCHRZ(1:NPA) = PREAL(3,:)
      ENDIF
      ! the base grid (avism.f90:366-382)
      LADOB = LADO0-(LADO0/NB)
      ALLOCATE(RADX(0:NB+1), RADY(0:NB+1), RADZ(0:NB+1))
      CALL MALLA(NB,NB,NB,LADOB)
      ALLOCATE(RADX0(0:NB+1), RADY0(0:NB+1), RADZ0(0:NB+1))
      RADX0 = RADX; RADY0 = RADY; RADZ0 = RADZ
      DX0 = DX; DY0 = DY; DZ0 = DZ
      DEALLOCATE(RADX, RADY, RADZ)
      ! the gas as READ_MASCLET leaves it: density contrast + 1 (reader.f90:392-398)
      ALLOCATE(U1G(NB,NB,NB), U2G(NB,NB,NB), U3G(NB,NB,NB), U4G(NB,NB,NB), U1S(NB,NB,NB), U1DM(NB,NB,NB))
      OPEN(2, FILE=TRIM(JOB)//'/amr_base.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
      READ(2) U1G, U2G, U3G, U4G
      CLOSE(2)
      U1S = -1.; U1DM = -1.
      U1G = U1G + 1.; U1S = U1S + 1.; U1DM = U1DM + 1.
      D1 = MAX(MAXVAL(PATCHNX), 1); D2 = MAX(MAXVAL(PATCHNY), 1); D3 = MAX(MAXVAL(PATCHNZ), 1)
      D4 = MAX(SUM(NPATCH(0:NL2)), 1)
      ALLOCATE(U11G(D1,D2,D3,D4), U12G(D1,D2,D3,D4), U13G(D1,D2,D3,D4), U14G(D1,D2,D3,D4))
      ALLOCATE(U11DM(D1,D2,D3,D4), U11S(D1,D2,D3,D4))
      U11G = -1.; U11DM = -1.; U11S = -1.
      U12G = 0.; U13G = 0.; U14G = 0.
      IF (NPA .GT. 0) THEN
         OPEN(2, FILE=TRIM(JOB)//'/amr_patch.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
         DO IP = 1, NPA
            N1 = PATCHNX(IP); N2 = PATCHNY(IP); N3 = PATCHNZ(IP)
            READ(2) U11G(1:N1,1:N2,1:N3,IP)
            READ(2) U12G(1:N1,1:N2,1:N3,IP)
            READ(2) U13G(1:N1,1:N2,1:N3,IP)
            READ(2) U14G(1:N1,1:N2,1:N3,IP)
         ENDDO
         CLOSE(2)
      ENDIF
      U11G = U11G + 1.; U11DM = U11DM + 1.; U11S = U11S + 1.
      ! the level's grid (avism.f90:750, 776-830)
      NXL = INT(REAL(NB)*(2.**(IRL)))
      LADOB = LADO0-(LADO0/NXL)
      ALLOCATE(RADX(0:NXL+1), RADY(0:NXL+1), RADZ(0:NXL+1))
      CALL MALLA(NXL,NXL,NXL,LADOB)
      DXL = DX
      RXL1 = RADX(1)
      ALLOCATE(U1GCO(NXL,NXL,NXL), U2GCO(NXL,NXL,NXL), U3GCO(NXL,NXL,NXL), U4GCO(NXL,NXL,NXL))
      ALLOCATE(U1SCO(NXL,NXL,NXL), U1DMCO(NXL,NXL,NXL), DIVERGCO(NXL,NXL,NXL), DIVERDMCO(NXL,NXL,NXL))
      ALLOCATE(DIVB(NXL,NXL,NXL))
      U1GCO = 0.; U2GCO = 0.; U3GCO = 0.; U4GCO = 0.; U1SCO = 0.; U1DMCO = 0.; DIVERGCO = 0.; DIVERDMCO = 0.
      IF (IRL .LT. 0) CALL SMOOTH_2(NXL,NXL,NXL)
      IF (IRL .EQ. 0) THEN
         U1GCO = U1G; U2GCO = U2G; U3GCO = U3G; U4GCO = U4G
      ENDIF
      IF (IRL .LE. 0) THEN
         CALL DIVER_UNIFORM(NXL,NXL,NXL,DXL,DXL,DXL,U2GCO,U3GCO,U4GCO,DIVB,FLAG_PERIOD)
         DIVERGCO = DIVB
      ELSE
         CALL DIVER_FINA_GAS(IRL)
         CALL VMESH(IRL, NXL, NXL, NXL, DXL, DXL, DXL, RXL1, RXL1, RXL1, NB, NB, NB)
         U1GCO = UGR
         DIVERGCO = DIVR
      ENDIF
      OPEN(3, FILE=TRIM(JOB)//'/out_amr.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) U1GCO, DIVERGCO
      CLOSE(3)
   END SUBROUTINE RUN_AMR

   ! READ_MASCLET of iteration ITER (amr.in: ITER, NPALEV, NLEVELS) with a base
   ! grid of NB^3 cells and room for NPMAX particles; NL2 = 1.  Writes
   ! out_masclet_head.bin (NDXYZ0, NL, NPARTT, NPATCH(1)), out_masclet_gas.bin
   ! (U1G, U2G, U3G, U4G as the reader leaves them), out_masclet_patch.bin (per
   ! level-1 patch PATCHNX..PATCHZ, PATCHRX..PATCHRZ, U11G..U14G) and
   ! out_masclet_part.bin (RXPA, RYPA, RZPA, U2PA, U3PA, U4PA, MASAP).
   SUBROUTINE RUN_READ_MASCLET(NB, NPMAX)
      INTEGER :: NB, NPMAX
      INTEGER :: ITER, NDXYZ0, NLR, IP, N1, N2, N3
      REAL*4 :: T, ZETAR
      OPEN(1, FILE=TRIM(JOB)//'/amr.in', STATUS='OLD', ACTION='READ')
      READ(1,*) ITER, NPALEV, NLEVELS
      CLOSE(1)
      NHYX = NB; NHYY = NB; NHYZ = NB
      PARTIRED = NPMAX
      NL2 = 1
      CALL READ_MASCLET(ITER, NDXYZ0, NLR, T, ZETAR)
      OPEN(3, FILE=TRIM(JOB)//'/out_masclet_head.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) NDXYZ0, NLR, INT(NPARTT, KIND=4), NPATCH(1)
      CLOSE(3)
      OPEN(3, FILE=TRIM(JOB)//'/out_masclet_gas.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) U1G, U2G, U3G, U4G
      CLOSE(3)
      OPEN(3, FILE=TRIM(JOB)//'/out_masclet_patch.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      DO IP = 1, NPATCH(1)
         N1 = PATCHNX(IP); N2 = PATCHNY(IP); N3 = PATCHNZ(IP)
         WRITE(3) N1, N2, N3, PATCHX(IP), PATCHY(IP), PATCHZ(IP), PATCHRX(IP), PATCHRY(IP), PATCHRZ(IP)
         WRITE(3) U11G(1:N1,1:N2,1:N3,IP), U12G(1:N1,1:N2,1:N3,IP), U13G(1:N1,1:N2,1:N3,IP), U14G(1:N1,1:N2,1:N3,IP)
      ENDDO
      CLOSE(3)
      OPEN(3, FILE=TRIM(JOB)//'/out_masclet_part.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
      WRITE(3) RXPA(1:NPARTT), RYPA(1:NPARTT), RZPA(1:NPARTT), U2PA(1:NPARTT), U3PA(1:NPARTT), U4PA(1:NPARTT), &
               MASAP(1:NPARTT)
      CLOSE(3)
   END SUBROUTINE RUN_READ_MASCLET

   SUBROUTINE DROP_VOID(ID)
      INTEGER :: ID
      WHERE (MARCA .EQ. ID)
         MARCA = 0
      END WHERE
      UVOID(ID) = 0
      WHERE (UVOID .EQ. ID)
         UVOID = 0
      END WHERE
      VOLNEW(ID) = 0.
      NCELLV(ID) = 0
   END SUBROUTINE DROP_VOID

   SUBROUTINE SET_VOLNEW()
      INTEGER :: K, ID
      VOLNEW = 0.
      DO K = 1, NVOID
         ID = INDICE(K)
         IF (UVOID(ID) .GE. 0) CYCLE
         VOLNEW(ID) = DBLE(NCELLV(ID))*DBLE(DXX*DXX*DXX)
      ENDDO
   END SUBROUTINE SET_VOLNEW

   SUBROUTINE RMIN_FILTER()
      INTEGER :: K, ID
      DO K = 1, NVOID
         ID = INDICE(K)
         IF (UVOID(ID) .GE. 0) CYCLE
         IF (VOLNEW(ID) .LT. 4./3.*PI*RMIN**3) CALL DROP_VOID(ID)
      ENDDO
   END SUBROUTINE RMIN_FILTER

END PROGRAM REF_HARNESS
