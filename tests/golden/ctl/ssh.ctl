DSET    ^ssh.dat                                                                                                                         
TITLE    SSH, m
UNDEF   -0.10000E+33  ! gap value
XDEF     285  LINEAR     27.525000          0.50000000E-01
YDEF     159  LINEAR     40.940000          0.40000000E-01
ZDEF       1  LINEAR     0.0000000           1.0000000    
TDEF       5  LINEAR  00:00Z01JAN1979        1dy
VARS 1  ! Number of variables
ssh            1  1 VARIABLES 
ENDVARS

