DSET    ^minutes.dat                                                                                                                     
TITLE    two-minute step
UNDEF   -0.10000E+33  ! gap value
XDEF       4  LINEAR     1.5000000           1.5000000    
YDEF       3  LINEAR     2.5000000           2.5000000    
ZDEF       1  LINEAR     0.0000000           1.0000000    
TDEF       1  LINEAR  00:00Z01JAN0005        1mn
VARS 1  ! Number of variables
x            1  1 VARIABLES 
ENDVARS

