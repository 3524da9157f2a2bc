DSET    ^year.dat                                                                                                                        
TITLE    
UNDEF   -0.99900E+03  ! gap value
XDEF      10  LINEAR     0.0000000          0.10000000E-02
YDEF      10  LINEAR     0.0000000           123456.00    
ZDEF       1  LINEAR     0.0000000           1.0000000    
TDEF       2  LINEAR  01:02Z04JUL1900        1yr
VARS 1  ! Number of variables
y            1  1 VARIABLES 
ENDVARS

