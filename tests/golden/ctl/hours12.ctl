DSET    ^hours12.dat                                                                                                                     
TITLE    barotropic state, 2048x1536
UNDEF   -0.10000E+33  ! gap value
XDEF    2048  LINEAR     26.643750          0.90000000E-02
YDEF    1536  LINEAR     40.145500          0.90000000E-02
ZDEF       1  LINEAR     0.0000000           1.0000000    
TDEF     100  LINEAR  00:00Z02DEC2026       12hr
VARS 1  ! Number of variables
ubrtr            1  1 VARIABLES 
ENDVARS

