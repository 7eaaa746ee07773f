C Record length of the .kta files for the REAL(8) build of the reference:
C every header and k-value record is 8 bytes wide.  This is the
C per-platform knob that radtran/rtm_util/isys.f sets to 4 for REAL(4).
      INTEGER FUNCTION ISYS()
      ISYS = 8
      RETURN
      END
