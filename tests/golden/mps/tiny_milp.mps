NAME          TINYMILP
OBJSENSE
 MIN
ROWS
 N  OBJ
 L  CAPACITY
COLUMNS
    MARK0000  'MARKER'                 'INTORG'
    X         OBJ      -2.0       CAPACITY 3.0
    Y         OBJ      -1.5       CAPACITY 2.0
    MARK0001  'MARKER'                 'INTEND'
RHS
    RHS1      CAPACITY  4.0
BOUNDS
 BV BND1      X
 BV BND1      Y
ENDATA
