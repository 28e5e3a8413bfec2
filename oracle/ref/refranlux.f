!     refranlux.f -- test harness around the reference's RANLUX
!     (f/luxury.f), built into oracle/_ref/refranlux by oracle/ref/Makefile.
!     Reads commands from stdin, one integer code per line followed by its
!     arguments; every result line starts with 'D' (the reference's own
!     messages on unit 6 are interleaved and ignored by the reader):
!       1 lux ins k1 k2   CALL RLUXGO(lux, ins, k1, k2)
!       2 n               CALL RANLUX(rvec, n); n bit patterns (TRANSFER)
!       3                 CALL RLUXUT(isd); 25 integers
!       4 isd(1..25)      CALL RLUXIN(isd)
!       5                 CALL RLUXAT(l, i, k1, k2); 4 integers
!       6 rank            the seeding of f/measinit.f:214-216; 16 integers
      PROGRAM refranlux
      INTEGER cmd, lux, ins, k1, k2, n, i, isd(25), l4(4), iseed(16)
      REAL rvec(200000), rv16(16)
   10 READ (5, *, END = 99) cmd
      IF (cmd .EQ. 1) THEN
        READ (5, *) lux, ins, k1, k2
        CALL rluxgo(lux, ins, k1, k2)
      ELSE IF (cmd .EQ. 2) THEN
        READ (5, *) n
        CALL ranlux(rvec, n)
        DO i = 1, n
          WRITE (6, '(A,I12)') 'D', TRANSFER(rvec(i), 0)
        END DO
      ELSE IF (cmd .EQ. 3) THEN
        CALL rluxut(isd)
        DO i = 1, 25
          WRITE (6, '(A,I12)') 'D', isd(i)
        END DO
      ELSE IF (cmd .EQ. 4) THEN
        READ (5, *) (isd(i), i = 1, 25)
        CALL rluxin(isd)
      ELSE IF (cmd .EQ. 5) THEN
        CALL rluxat(l4(1), l4(2), l4(3), l4(4))
        DO i = 1, 4
          WRITE (6, '(A,I12)') 'D', l4(i)
        END DO
      ELSE IF (cmd .EQ. 6) THEN
        READ (5, *) n
        CALL rluxgo(3, 314159 + n, 0, 0)
        CALL ranlux(rv16, 16)
        iseed = int(10**6 * rv16)
        DO i = 1, 16
          WRITE (6, '(A,I12)') 'D', iseed(i)
        END DO
      END IF
      GOTO 10
   99 STOP
      END
