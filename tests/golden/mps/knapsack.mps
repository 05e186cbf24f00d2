NAME          KNAP
ROWS
 N  OBJ
 L  CAP
COLUMNS
    MARKER                 'MARKER'                 'INTORG'
    A         OBJ       -10.0      CAP       3.0
    B         OBJ       -13.0      CAP       5.0
    C         OBJ       -7.0       CAP       2.0
    D         OBJ       -9.0       CAP       4.0
    MARKER                 'MARKER'                 'INTEND'
RHS
    RHS       CAP       7.0
BOUNDS
 BV BND       A
 BV BND       B
 BV BND       C
 BV BND       D
ENDATA
