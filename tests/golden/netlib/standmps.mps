NAME          STANDMPS (STANDATA) SIZE: N=1075, M=468, NZ=4754
ROWS
 E  A..1..S1
 E  A..2..S1
 E  A..3..S1
 E  A..4..S1
 E  A..5..S1
 E  A..6..S1
 E  A..1..S2
 E  A..2..S2
 E  A..3..S2
 E  A..4..S2
 E  A..5..S2
 E  A..6..S2
 E  BP11..S1
 E  BP21..S1
 E  BP12..S1
 E  BP22..S1
 E  BP13..S1
 E  BP23..S1
 E  BP14..S1
 E  BP24..S1
 E  BP15..S1
 E  BP25..S1
 E  BP16..S1
 E  BP26..S1
 E  BP11..S2
 E  BP21..S2
 E  BP12..S2
 E  BP22..S2
 E  BP13..S2
 E  BP23..S2
 E  BP14..S2
 E  BP24..S2
 E  BP15..S2
 E  BP25..S2
 E  BP16..S2
 E  BP26..S2
 E  BP11..T1
 E  BP21..T1
 E  BP12..T1
 E  BP22..T1
 E  BP13..T1
 E  BP23..T1
 E  BP14..T1
 E  BP24..T1
 E  BP15..T1
 E  BP25..T1
 E  BP16..T1
 E  BP26..T1
 E  BP11..T2
 E  BP21..T2
 E  BP12..T2
 E  BP22..T2
 E  BP13..T2
 E  BP23..T2
 E  BP14..T2
 E  BP24..T2
 E  BP15..T2
 E  BP25..T2
 E  BP16..T2
 E  BP26..T2
 E  BP11..T3
 E  BP21..T3
 E  BP12..T3
 E  BP22..T3
 E  BP13..T3
 E  BP23..T3
 E  BP14..T3
 E  BP24..T3
 E  BP15..T3
 E  BP25..T3
 E  BP16..T3
 E  BP26..T3
 E  BP11..T4
 E  BP21..T4
 E  BP12..T4
 E  BP22..T4
 E  BP13..T4
 E  BP23..T4
 E  BP14..T4
 E  BP24..T4
 E  BP15..T4
 E  BP25..T4
 E  BP16..T4
 E  BP26..T4
 E  C..1..T1
 E  C..2..T1
 E  C..3..T1
 E  C..4..T1
 E  C..5..T1
 E  C..6..T1
 E  C..1..T2
 E  C..2..T2
 E  C..3..T2
 E  C..4..T2
 E  C..5..T2
 E  C..6..T2
 E  C..1..T3
 E  C..2..T3
 E  C..3..T3
 E  C..4..T3
 E  C..5..T3
 E  C..6..T3
 E  C..1..T4
 E  C..2..T4
 E  C..3..T4
 E  C..4..T4
 E  C..5..T4
 E  C..6..T4
 E  FCC.....
 E  FIC.....
 E  FOT.....
 E  FTR.....
 N  FAT...J.
 L  FIN...1.
 L  FIN...2.
 L  FIN...3.
 L  I..1..T1
 L  I..2..T1
 L  I..3..T1
 L  I..4..T1
 L  I..5..T1
 L  I..6..T1
 L  I..1..T2
 L  I..2..T2
 L  I..3..T2
 L  I..4..T2
 L  I..5..T2
 L  I..6..T2
 L  I..1..T3
 L  I..2..T3
 L  I..3..T3
 L  I..4..T3
 L  I..5..T3
 L  I..6..T3
 L  I..1..T4
 L  I..2..T4
 L  I..3..T4
 L  I..4..T4
 L  I..5..T4
 L  I..6..T4
 L  KM.1..T1
 L  KN.1..T1
 L  KM.3..T1
 L  KN.3..T1
 L  KM.5..T1
 L  KN.5..T1
 L  KM.1..T2
 L  KN.1..T2
 L  KM.3..T2
 L  KN.3..T2
 L  KM.5..T2
 L  KN.5..T2
 L  KM.1..T3
 L  KN.1..T3
 L  KM.3..T3
 L  KN.3..T3
 L  KM.5..T3
 L  KN.5..T3
 L  KM.1..T4
 L  KN.1..T4
 L  KM.3..T4
 L  KN.3..T4
 L  KM.5..T4
 L  KN.5..T4
 L  LY....T1
 L  LC.1..T1
 L  LM.1..T1
 L  LN.1..T1
 L  LC.3..T1
 L  LM.3..T1
 L  LN.3..T1
 L  LC.5..T1
 L  LM.5..T1
 L  LN.5..T1
 L  LY....T2
 L  LC.1..T2
 L  LM.1..T2
 L  LN.1..T2
 L  LC.3..T2
 L  LM.3..T2
 L  LN.3..T2
 L  LC.5..T2
 L  LM.5..T2
 L  LN.5..T2
 L  LY....T3
 L  LC.1..T3
 L  LM.1..T3
 L  LN.1..T3
 L  LC.3..T3
 L  LM.3..T3
 L  LN.3..T3
 L  LC.5..T3
 L  LM.5..T3
 L  LN.5..T3
 L  LY....T4
 L  LC.1..T4
 L  LM.1..T4
 L  LN.1..T4
 L  LC.3..T4
 L  LM.3..T4
 L  LN.3..T4
 L  LC.5..T4
 L  LM.5..T4
 L  LN.5..T4
 L  TYEAR1T1
 L  TYEAR2T1
 L  TYEAR3T1
 E  TB.1S1T1
 E  TB.1S2T1
 L  TM.1S1T1
 L  TM.1S2T1
 L  TN.1S1T1
 L  TN.1S2T1
 E  TB.2S1T1
 E  TB.2S2T1
 L  TM.2S1T1
 L  TM.2S2T1
 L  TN.2S1T1
 L  TN.2S2T1
 E  TB.3S1T1
 E  TB.3S2T1
 L  TM.3S1T1
 L  TM.3S2T1
 L  TN.3S1T1
 L  TN.3S2T1
 E  TB.4S1T1
 E  TB.4S2T1
 L  TM.4S1T1
 L  TM.4S2T1
 L  TN.4S1T1
 L  TN.4S2T1
 E  TB.5S1T1
 E  TB.5S2T1
 L  TM.5S1T1
 L  TM.5S2T1
 L  TN.5S1T1
 L  TN.5S2T1
 E  TB.6S1T1
 E  TB.6S2T1
 L  TM.6S1T1
 L  TM.6S2T1
 L  TN.6S1T1
 L  TN.6S2T1
 L  TYEAR1T2
 L  TYEAR2T2
 L  TYEAR3T2
 E  TB.1S1T2
 E  TB.1S2T2
 L  TM.1S1T2
 L  TM.1S2T2
 L  TN.1S1T2
 L  TN.1S2T2
 E  TB.2S1T2
 E  TB.2S2T2
 L  TM.2S1T2
 L  TM.2S2T2
 L  TN.2S1T2
 L  TN.2S2T2
 E  TB.3S1T2
 E  TB.3S2T2
 L  TM.3S1T2
 L  TM.3S2T2
 L  TN.3S1T2
 L  TN.3S2T2
 E  TB.4S1T2
 E  TB.4S2T2
 L  TM.4S1T2
 L  TM.4S2T2
 L  TN.4S1T2
 L  TN.4S2T2
 E  TB.5S1T2
 E  TB.5S2T2
 L  TM.5S1T2
 L  TM.5S2T2
 L  TN.5S1T2
 L  TN.5S2T2
 E  TB.6S1T2
 E  TB.6S2T2
 L  TM.6S1T2
 L  TM.6S2T2
 L  TN.6S1T2
 L  TN.6S2T2
 L  TYEAR1T3
 L  TYEAR2T3
 L  TYEAR3T3
 E  TB.1S1T3
 E  TB.1S2T3
 L  TM.1S1T3
 L  TM.1S2T3
 L  TN.1S1T3
 L  TN.1S2T3
 E  TB.2S1T3
 E  TB.2S2T3
 L  TM.2S1T3
 L  TM.2S2T3
 L  TN.2S1T3
 L  TN.2S2T3
 E  TB.3S1T3
 E  TB.3S2T3
 L  TM.3S1T3
 L  TM.3S2T3
 L  TN.3S1T3
 L  TN.3S2T3
 E  TB.4S1T3
 E  TB.4S2T3
 L  TM.4S1T3
 L  TM.4S2T3
 L  TN.4S1T3
 L  TN.4S2T3
 E  TB.5S1T3
 E  TB.5S2T3
 L  TM.5S1T3
 L  TM.5S2T3
 L  TN.5S1T3
 L  TN.5S2T3
 E  TB.6S1T3
 E  TB.6S2T3
 L  TM.6S1T3
 L  TM.6S2T3
 L  TN.6S1T3
 L  TN.6S2T3
 L  TYEAR1T4
 L  TYEAR2T4
 L  TYEAR3T4
 E  TB.1S1T4
 E  TB.1S2T4
 L  TM.1S1T4
 L  TM.1S2T4
 L  TN.1S1T4
 L  TN.1S2T4
 E  TB.2S1T4
 E  TB.2S2T4
 L  TM.2S1T4
 L  TM.2S2T4
 L  TN.2S1T4
 L  TN.2S2T4
 E  TB.3S1T4
 E  TB.3S2T4
 L  TM.3S1T4
 L  TM.3S2T4
 L  TN.3S1T4
 L  TN.3S2T4
 E  TB.4S1T4
 E  TB.4S2T4
 L  TM.4S1T4
 L  TM.4S2T4
 L  TN.4S1T4
 L  TN.4S2T4
 E  TB.5S1T4
 E  TB.5S2T4
 L  TM.5S1T4
 L  TM.5S2T4
 L  TN.5S1T4
 L  TN.5S2T4
 E  TB.6S1T4
 E  TB.6S2T4
 L  TM.6S1T4
 L  TM.6S2T4
 L  TN.6S1T4
 L  TN.6S2T4
 E  ZP11..01
 E  ZP21..01
 E  ZP12..01
 E  ZP22..01
 E  ZP13..01
 E  ZP23..01
 E  ZP14..01
 E  ZP24..01
 E  ZP15..01
 E  ZP25..01
 E  ZP16..01
 E  ZP26..01
 E  ZP11..02
 E  ZP21..02
 E  ZP12..02
 E  ZP22..02
 E  ZP13..02
 E  ZP23..02
 E  ZP14..02
 E  ZP24..02
 E  ZP15..02
 E  ZP25..02
 E  ZP16..02
 E  ZP26..02
 E  ZP11..03
 E  ZP21..03
 E  ZP12..03
 E  ZP22..03
 E  ZP13..03
 E  ZP23..03
 E  ZP14..03
 E  ZP24..03
 E  ZP15..03
 E  ZP25..03
 E  ZP16..03
 E  ZP26..03
 E  ZP11..04
 E  ZP21..04
 E  ZP12..04
 E  ZP22..04
 E  ZP13..04
 E  ZP23..04
 E  ZP14..04
 E  ZP24..04
 E  ZP15..04
 E  ZP25..04
 E  ZP16..04
 E  ZP26..04
 E  ZP11..05
 E  ZP21..05
 E  ZP12..05
 E  ZP22..05
 E  ZP13..05
 E  ZP23..05
 E  ZP14..05
 E  ZP24..05
 E  ZP15..05
 E  ZP25..05
 E  ZP16..05
 E  ZP26..05
 E  ZP11..06
 E  ZP21..06
 E  ZP12..06
 E  ZP22..06
 E  ZP13..06
 E  ZP23..06
 E  ZP14..06
 E  ZP24..06
 E  ZP15..06
 E  ZP25..06
 E  ZP16..06
 E  ZP26..06
 E  ZP11..07
 E  ZP21..07
 E  ZP12..07
 E  ZP22..07
 E  ZP13..07
 E  ZP23..07
 E  ZP14..07
 E  ZP24..07
 E  ZP15..07
 E  ZP25..07
 E  ZP16..07
 E  ZP26..07
 E  ZP11..08
 E  ZP21..08
 E  ZP12..08
 E  ZP22..08
 E  ZP13..08
 E  ZP23..08
 E  ZP14..08
 E  ZP24..08
 E  ZP15..08
 E  ZP25..08
 E  ZP16..08
 E  ZP26..08
 E  ZP11..09
 E  ZP21..09
 E  ZP12..09
 E  ZP22..09
 E  ZP13..09
 E  ZP23..09
 E  ZP14..09
 E  ZP24..09
 E  ZP15..09
 E  ZP25..09
 E  ZP16..09
 E  ZP26..09
COLUMNS
    A..1X1S1  A..1..S1           -1.
    A..1X2S1  A..1..S1           -1.
    A..1X3S1  A..1..S1           -1.
    AP11..S1  BP11..S1            1.   A..1..S1            1.
    AP21..S1  BP21..S1            1.   A..1..S1            1.
    A..2X1S1  A..2..S1           -1.
    A..2X2S1  A..2..S1           -1.   FOT.....           2.5
    A..2X3S1  A..2..S1           -1.   FOT.....            3.
    AP12..S1  BP12..S1            1.   A..2..S1            1.
    AP22..S1  BP22..S1            1.   A..2..S1            1.
    A..3X1S1  A..3..S1           -1.
    A..3X2S1  A..3..S1           -1.
    A..3X3S1  A..3..S1           -1.
    AP13..S1  BP13..S1            1.   A..3..S1            1.
    AP23..S1  BP23..S1            1.   A..3..S1            1.
    A..4X1S1  A..4..S1           -1.
    A..4X2S1  A..4..S1           -1.   FOT.....           2.5
    A..4X3S1  A..4..S1           -1.   FOT.....            3.
    AP14..S1  BP14..S1            1.   A..4..S1            1.
    AP24..S1  BP24..S1            1.   A..4..S1            1.
    A..5X1S1  A..5..S1           -1.
    A..5X2S1  A..5..S1           -1.
    A..5X3S1  A..5..S1           -1.
    AP15..S1  BP15..S1            1.   A..5..S1            1.
    AP25..S1  BP25..S1            1.   A..5..S1            1.
    A..6X1S1  A..6..S1           -1.
    A..6X2S1  A..6..S1           -1.   FOT.....           2.5
    A..6X3S1  A..6..S1           -1.   FOT.....            3.
    AP16..S1  BP16..S1            1.   A..6..S1            1.
    AP26..S1  BP26..S1            1.   A..6..S1            1.
    A..1X1S2  A..1..S2           -1.
    A..1X2S2  A..1..S2           -1.
    A..1X3S2  A..1..S2           -1.
    AP11..S2  BP11..S2            1.   A..1..S2            1.
    AP21..S2  BP21..S2            1.   A..1..S2            1.
    A..2X1S2  A..2..S2           -1.
    A..2X2S2  A..2..S2           -1.   FOT.....           2.5
    A..2X3S2  A..2..S2           -1.   FOT.....            3.
    AP12..S2  BP12..S2            1.   A..2..S2            1.
    AP22..S2  BP22..S2            1.   A..2..S2            1.
    A..3X1S2  A..3..S2           -1.
    A..3X2S2  A..3..S2           -1.
    A..3X3S2  A..3..S2           -1.
    AP13..S2  BP13..S2            1.   A..3..S2            1.
    AP23..S2  BP23..S2            1.   A..3..S2            1.
    A..4X1S2  A..4..S2           -1.
    A..4X2S2  A..4..S2           -1.   FOT.....           2.5
    A..4X3S2  A..4..S2           -1.   FOT.....            3.
    AP14..S2  BP14..S2            1.   A..4..S2            1.
    AP24..S2  BP24..S2            1.   A..4..S2            1.
    A..5X1S2  A..5..S2           -1.
    A..5X2S2  A..5..S2           -1.
    A..5X3S2  A..5..S2           -1.
    AP15..S2  BP15..S2            1.   A..5..S2            1.
    AP25..S2  BP25..S2            1.   A..5..S2            1.
    A..6X1S2  A..6..S2           -1.
    A..6X2S2  A..6..S2           -1.   FOT.....           2.5
    A..6X3S2  A..6..S2           -1.   FOT.....            3.
    AP16..S2  BP16..S2            1.   A..6..S2            1.
    AP26..S2  BP26..S2            1.   A..6..S2            1.
    C..1..T1  C..1..T1           -1.   I..1..T1           -1.
    C..1..T1  C..2..T1            1.
    CIN1..T1  C..1..T1            1.   KM.1..T1            1.
    CIN1..T1  KN.1..T1           -1.   FCC.....       7.20725
    CIN1..T1  FIN...1.         31.75
    CVA1..T1  C..1..T1            1.   LM.1..T1            1.
    CVA1..T1  LN.1..T1           -1.   FCC.....        6.2652
    CVA1..T1  FIN...1.          27.6
    C..2..T1  C..2..T1           -1.   I..2..T1           -1.
    C..2..T1  C..3..T1            1.
    C..3..T1  C..3..T1           -1.   I..3..T1           -1.
    C..3..T1  C..4..T1            1.
    CIN3..T1  C..3..T1            1.   KM.3..T1            1.
    CIN3..T1  KN.3..T1           -1.   FCC.....          5.08
    CIN3..T1  FIN...2.         31.75
    CVA3..T1  C..3..T1            1.   LM.3..T1            1.
    CVA3..T1  LN.3..T1           -1.   FCC.....         4.416
    CVA3..T1  FIN...2.          27.6
    C..4..T1  C..4..T1           -1.   I..4..T1           -1.
    C..4..T1  C..5..T1            1.
    C..5..T1  C..5..T1           -1.   I..5..T1           -1.
    C..5..T1  C..6..T1            1.
    CIN5..T1  C..5..T1            1.   KM.5..T1            1.
    CIN5..T1  KN.5..T1           -1.   FCC.....       2.69875
    CIN5..T1  FIN...3.         31.75
    CVA5..T1  C..5..T1            1.   LM.5..T1            1.
    CVA5..T1  LN.5..T1           -1.   FCC.....         2.346
    CVA5..T1  FIN...3.          27.6
    C..6..T1  C..6..T1           -1.   I..6..T1           -1.
    C..1..T2  C..1..T2           -1.   I..1..T2           -1.
    C..1..T2  C..2..T2            1.
    CIN1..T2  C..1..T2            1.   KM.1..T2            1.
    CIN1..T2  KN.1..T2           -1.   FCC.....       7.20725
    CIN1..T2  FIN...1.         31.75
    CVA1..T2  C..1..T2            1.   LM.1..T2            1.
    CVA1..T2  LN.1..T2           -1.   FCC.....        6.2652
    CVA1..T2  FIN...1.          27.6
    C..2..T2  C..2..T2           -1.   I..2..T2           -1.
    C..2..T2  C..3..T2            1.
    C..3..T2  C..3..T2           -1.   I..3..T2           -1.
    C..3..T2  C..4..T2            1.
    CIN3..T2  C..3..T2            1.   KM.3..T2            1.
    CIN3..T2  KN.3..T2           -1.   FCC.....          5.08
    CIN3..T2  FIN...2.         31.75
    CVA3..T2  C..3..T2            1.   LM.3..T2            1.
    CVA3..T2  LN.3..T2           -1.   FCC.....         4.416
    CVA3..T2  FIN...2.          27.6
    C..4..T2  C..4..T2           -1.   I..4..T2           -1.
    C..4..T2  C..5..T2            1.
    C..5..T2  C..5..T2           -1.   I..5..T2           -1.
    C..5..T2  C..6..T2            1.
    CIN5..T2  C..5..T2            1.   KM.5..T2            1.
    CIN5..T2  KN.5..T2           -1.   FCC.....       2.69875
    CIN5..T2  FIN...3.         31.75
    CVA5..T2  C..5..T2            1.   LM.5..T2            1.
    CVA5..T2  LN.5..T2           -1.   FCC.....         2.346
    CVA5..T2  FIN...3.          27.6
    C..6..T2  C..6..T2           -1.   I..6..T2           -1.
    C..1..T3  C..1..T3           -1.   I..1..T3           -1.
    C..1..T3  C..2..T3            1.
    CIN1..T3  C..1..T3            1.   KM.1..T3            1.
    CIN1..T3  KN.1..T3           -1.   FCC.....       7.20725
    CIN1..T3  FIN...1.         31.75
    CVA1..T3  C..1..T3            1.   LM.1..T3            1.
    CVA1..T3  LN.1..T3           -1.   FCC.....        6.2652
    CVA1..T3  FIN...1.          27.6
    C..2..T3  C..2..T3           -1.   I..2..T3           -1.
    C..2..T3  C..3..T3            1.
    C..3..T3  C..3..T3           -1.   I..3..T3           -1.
    C..3..T3  C..4..T3            1.
    CIN3..T3  C..3..T3            1.   KM.3..T3            1.
    CIN3..T3  KN.3..T3           -1.   FCC.....          5.08
    CIN3..T3  FIN...2.         31.75
    CVA3..T3  C..3..T3            1.   LM.3..T3            1.
    CVA3..T3  LN.3..T3           -1.   FCC.....         4.416
    CVA3..T3  FIN...2.          27.6
    C..4..T3  C..4..T3           -1.   I..4..T3           -1.
    C..4..T3  C..5..T3            1.
    C..5..T3  C..5..T3           -1.   I..5..T3           -1.
    C..5..T3  C..6..T3            1.
    CIN5..T3  C..5..T3            1.   KM.5..T3            1.
    CIN5..T3  KN.5..T3           -1.   FCC.....       2.69875
    CIN5..T3  FIN...3.         31.75
    CVA5..T3  C..5..T3            1.   LM.5..T3            1.
    CVA5..T3  LN.5..T3           -1.   FCC.....         2.346
    CVA5..T3  FIN...3.          27.6
    C..6..T3  C..6..T3           -1.   I..6..T3           -1.
    C..1..T4  C..1..T4           -1.   I..1..T4           -1.
    C..1..T4  C..2..T4            1.
    CIN1..T4  C..1..T4            1.   KM.1..T4            1.
    CIN1..T4  KN.1..T4           -1.   FCC.....       7.20725
    CIN1..T4  FIN...1.         31.75
    CVA1..T4  C..1..T4            1.   LM.1..T4            1.
    CVA1..T4  LN.1..T4           -1.   FCC.....        6.2652
    CVA1..T4  FIN...1.          27.6
    C..2..T4  C..2..T4           -1.   I..2..T4           -1.
    C..2..T4  C..3..T4            1.
    C..3..T4  C..3..T4           -1.   I..3..T4           -1.
    C..3..T4  C..4..T4            1.
    CIN3..T4  C..3..T4            1.   KM.3..T4            1.
    CIN3..T4  KN.3..T4           -1.   FCC.....          5.08
    CIN3..T4  FIN...2.         31.75
    CVA3..T4  C..3..T4            1.   LM.3..T4            1.
    CVA3..T4  LN.3..T4           -1.   FCC.....         4.416
    CVA3..T4  FIN...2.          27.6
    C..4..T4  C..4..T4           -1.   I..4..T4           -1.
    C..4..T4  C..5..T4            1.
    C..5..T4  C..5..T4           -1.   I..5..T4           -1.
    C..5..T4  C..6..T4            1.
    CIN5..T4  C..5..T4            1.   KM.5..T4            1.
    CIN5..T4  KN.5..T4           -1.   FCC.....       2.69875
    CIN5..T4  FIN...3.         31.75
    CVA5..T4  C..5..T4            1.   LM.5..T4            1.
    CVA5..T4  LN.5..T4           -1.   FCC.....         2.346
    CVA5..T4  FIN...3.          27.6
    C..6..T4  C..6..T4           -1.   I..6..T4           -1.
    FCC.....  FCC.....           -1.   FAT...J.            1.
    FIC.....  FIC.....           -1.   FAT...J.            1.
    FOT.....  FOT.....           -1.   FAT...J.            1.
    FTR.....  FTR.....           -1.   FAT...J.            1.
    IP11..T1  BP11..T1           -1.   I..1..T1            1.
    IP11..T1  FIC.....            3.   BP12..T1            1.
    IP21..T1  BP21..T1           -1.   I..1..T1            1.
    IP21..T1  FIC.....            3.   BP22..T1            1.
    IP12..T1  BP12..T1           -1.   I..2..T1            1.
    IP12..T1  FIC.....            3.   BP13..T1            1.
    IP22..T1  BP22..T1           -1.   I..2..T1            1.
    IP22..T1  FIC.....            3.   BP23..T1            1.
    IP13..T1  BP13..T1           -1.   I..3..T1            1.
    IP13..T1  FIC.....            3.   BP14..T1            1.
    IP23..T1  BP23..T1           -1.   I..3..T1            1.
    IP23..T1  FIC.....            3.   BP24..T1            1.
    IP14..T1  BP14..T1           -1.   I..4..T1            1.
    IP14..T1  FIC.....            3.   BP15..T1            1.
    IP24..T1  BP24..T1           -1.   I..4..T1            1.
    IP24..T1  FIC.....            3.   BP25..T1            1.
    IP15..T1  BP15..T1           -1.   I..5..T1            1.
    IP15..T1  FIC.....            3.   BP16..T1            1.
    IP25..T1  BP25..T1           -1.   I..5..T1            1.
    IP25..T1  FIC.....            3.   BP26..T1            1.
    IP16..T1  BP16..T1           -1.   I..6..T1            1.
    IP16..T1  FIC.....            3.
    IP26..T1  BP26..T1           -1.   I..6..T1            1.
    IP26..T1  FIC.....            3.
    IP11..T2  BP11..T2           -1.   I..1..T2            1.
    IP11..T2  FIC.....            3.   BP12..T2            1.
    IP21..T2  BP21..T2           -1.   I..1..T2            1.
    IP21..T2  FIC.....            3.   BP22..T2            1.
    IP12..T2  BP12..T2           -1.   I..2..T2            1.
    IP12..T2  FIC.....            3.   BP13..T2            1.
    IP22..T2  BP22..T2           -1.   I..2..T2            1.
    IP22..T2  FIC.....            3.   BP23..T2            1.
    IP13..T2  BP13..T2           -1.   I..3..T2            1.
    IP13..T2  FIC.....            3.   BP14..T2            1.
    IP23..T2  BP23..T2           -1.   I..3..T2            1.
    IP23..T2  FIC.....            3.   BP24..T2            1.
    IP14..T2  BP14..T2           -1.   I..4..T2            1.
    IP14..T2  FIC.....            3.   BP15..T2            1.
    IP24..T2  BP24..T2           -1.   I..4..T2            1.
    IP24..T2  FIC.....            3.   BP25..T2            1.
    IP15..T2  BP15..T2           -1.   I..5..T2            1.
    IP15..T2  FIC.....            3.   BP16..T2            1.
    IP25..T2  BP25..T2           -1.   I..5..T2            1.
    IP25..T2  FIC.....            3.   BP26..T2            1.
    IP16..T2  BP16..T2           -1.   I..6..T2            1.
    IP16..T2  FIC.....            3.
    IP26..T2  BP26..T2           -1.   I..6..T2            1.
    IP26..T2  FIC.....            3.
    IP11..T3  BP11..T3           -1.   I..1..T3            1.
    IP11..T3  FIC.....            3.   BP12..T3            1.
    IP21..T3  BP21..T3           -1.   I..1..T3            1.
    IP21..T3  FIC.....            3.   BP22..T3            1.
    IP12..T3  BP12..T3           -1.   I..2..T3            1.
    IP12..T3  FIC.....            3.   BP13..T3            1.
    IP22..T3  BP22..T3           -1.   I..2..T3            1.
    IP22..T3  FIC.....            3.   BP23..T3            1.
    IP13..T3  BP13..T3           -1.   I..3..T3            1.
    IP13..T3  FIC.....            3.   BP14..T3            1.
    IP23..T3  BP23..T3           -1.   I..3..T3            1.
    IP23..T3  FIC.....            3.   BP24..T3            1.
    IP14..T3  BP14..T3           -1.   I..4..T3            1.
    IP14..T3  FIC.....            3.   BP15..T3            1.
    IP24..T3  BP24..T3           -1.   I..4..T3            1.
    IP24..T3  FIC.....            3.   BP25..T3            1.
    IP15..T3  BP15..T3           -1.   I..5..T3            1.
    IP15..T3  FIC.....            3.   BP16..T3            1.
    IP25..T3  BP25..T3           -1.   I..5..T3            1.
    IP25..T3  FIC.....            3.   BP26..T3            1.
    IP16..T3  BP16..T3           -1.   I..6..T3            1.
    IP16..T3  FIC.....            3.
    IP26..T3  BP26..T3           -1.   I..6..T3            1.
    IP26..T3  FIC.....            3.
    IP11..T4  BP11..T4           -1.   I..1..T4            1.
    IP11..T4  FIC.....            3.   BP12..T4            1.
    IP21..T4  BP21..T4           -1.   I..1..T4            1.
    IP21..T4  FIC.....            3.   BP22..T4            1.
    IP12..T4  BP12..T4           -1.   I..2..T4            1.
    IP12..T4  FIC.....            3.   BP13..T4            1.
    IP22..T4  BP22..T4           -1.   I..2..T4            1.
    IP22..T4  FIC.....            3.   BP23..T4            1.
    IP13..T4  BP13..T4           -1.   I..3..T4            1.
    IP13..T4  FIC.....            3.   BP14..T4            1.
    IP23..T4  BP23..T4           -1.   I..3..T4            1.
    IP23..T4  FIC.....            3.   BP24..T4            1.
    IP14..T4  BP14..T4           -1.   I..4..T4            1.
    IP14..T4  FIC.....            3.   BP15..T4            1.
    IP24..T4  BP24..T4           -1.   I..4..T4            1.
    IP24..T4  FIC.....            3.   BP25..T4            1.
    IP15..T4  BP15..T4           -1.   I..5..T4            1.
    IP15..T4  FIC.....            3.   BP16..T4            1.
    IP25..T4  BP25..T4           -1.   I..5..T4            1.
    IP25..T4  FIC.....            3.   BP26..T4            1.
    IP16..T4  BP16..T4           -1.   I..6..T4            1.
    IP16..T4  FIC.....            3.
    IP26..T4  BP26..T4           -1.   I..6..T4            1.
    IP26..T4  FIC.....            3.
    KBU1S1T1  TN.1S1T1           10.   TM.1S1T1          -25.
    KBU1S2T1  TN.1S2T1           10.   TM.1S2T1          -25.
    KIN1..T1  LY....T1            1.   KM.1..T1          -15.
    KIN1..T1  KN.1..T1            3.   FCC.....          68.1
    KIN1..T1  FIN...1.          300.   TYEAR1T1          -25.
    KIN1..T1  TYEAR2T1          -25.   LC.3..T1           -1.
    KIN1..T1  TYEAR3T1          -25.   LC.5..T1           -1.
    KVA1..T1  LM.1..T1          -10.   LN.1..T1            2.
    KVA1..T1  FCC.....        24.289   FIN...1.          107.
    KVA1..T1  LC.1..T1            1.
    KBU2S1T1  TN.2S1T1           10.   TM.2S1T1          -25.
    KBU2S2T1  TN.2S2T1           10.   TM.2S2T1          -25.
    KBU3S1T1  TN.3S1T1           10.   TM.3S1T1          -25.
    KBU3S2T1  TN.3S2T1           10.   TM.3S2T1          -25.
    KIN3..T1  LY....T1            1.   KM.3..T1          -15.
    KIN3..T1  KN.3..T1            3.   FCC.....           48.
    KIN3..T1  FIN...2.          300.   TYEAR2T1          -25.
    KIN3..T1  TYEAR3T1          -25.   LC.5..T1           -1.
    KVA3..T1  LM.3..T1          -10.   LN.3..T1            2.
    KVA3..T1  FCC.....         17.12   FIN...2.          107.
    KVA3..T1  LC.3..T1            1.
    KBU4S1T1  TN.4S1T1           10.   TM.4S1T1          -25.
    KBU4S2T1  TN.4S2T1           10.   TM.4S2T1          -25.
    KBU5S1T1  TN.5S1T1           10.   TM.5S1T1          -25.
    KBU5S2T1  TN.5S2T1           10.   TM.5S2T1          -25.
    KIN5..T1  LY....T1            1.   KM.5..T1          -15.
    KIN5..T1  KN.5..T1            3.   FCC.....          25.5
    KIN5..T1  FIN...3.          300.   TYEAR3T1          -25.
    KVA5..T1  LM.5..T1          -10.   LN.5..T1            2.
    KVA5..T1  FCC.....         9.095   FIN...3.          107.
    KVA5..T1  LC.5..T1            1.
    KBU6S1T1  TN.6S1T1           10.   TM.6S1T1          -25.
    KBU6S2T1  TN.6S2T1           10.   TM.6S2T1          -25.
    KBU1S1T2  TN.1S1T2           10.   TM.1S1T2          -25.
    KBU1S2T2  TN.1S2T2           10.   TM.1S2T2          -25.
    KIN1..T2  LY....T2            1.   KM.1..T2          -15.
    KIN1..T2  KN.1..T2            3.   FCC.....          68.1
    KIN1..T2  FIN...1.          300.   TYEAR1T2          -25.
    KIN1..T2  TYEAR2T2          -25.   LC.3..T2           -1.
    KIN1..T2  TYEAR3T2          -25.   LC.5..T2           -1.
    KVA1..T2  LM.1..T2          -10.   LN.1..T2            2.
    KVA1..T2  FCC.....        24.289   FIN...1.          107.
    KVA1..T2  LC.1..T2            1.
    KBU2S1T2  TN.2S1T2           10.   TM.2S1T2          -25.
    KBU2S2T2  TN.2S2T2           10.   TM.2S2T2          -25.
    KBU3S1T2  TN.3S1T2           10.   TM.3S1T2          -25.
    KBU3S2T2  TN.3S2T2           10.   TM.3S2T2          -25.
    KIN3..T2  LY....T2            1.   KM.3..T2          -15.
    KIN3..T2  KN.3..T2            3.   FCC.....           48.
    KIN3..T2  FIN...2.          300.   TYEAR2T2          -25.
    KIN3..T2  TYEAR3T2          -25.   LC.5..T2           -1.
    KVA3..T2  LM.3..T2          -10.   LN.3..T2            2.
    KVA3..T2  FCC.....         17.12   FIN...2.          107.
    KVA3..T2  LC.3..T2            1.
    KBU4S1T2  TN.4S1T2           10.   TM.4S1T2          -25.
    KBU4S2T2  TN.4S2T2           10.   TM.4S2T2          -25.
    KBU5S1T2  TN.5S1T2           10.   TM.5S1T2          -25.
    KBU5S2T2  TN.5S2T2           10.   TM.5S2T2          -25.
    KIN5..T2  LY....T2            1.   KM.5..T2          -15.
    KIN5..T2  KN.5..T2            3.   FCC.....          25.5
    KIN5..T2  FIN...3.          300.   TYEAR3T2          -25.
    KVA5..T2  LM.5..T2          -10.   LN.5..T2            2.
    KVA5..T2  FCC.....         9.095   FIN...3.          107.
    KVA5..T2  LC.5..T2            1.
    KBU6S1T2  TN.6S1T2           10.   TM.6S1T2          -25.
    KBU6S2T2  TN.6S2T2           10.   TM.6S2T2          -25.
    KBU1S1T3  TN.1S1T3           10.   TM.1S1T3          -25.
    KBU1S2T3  TN.1S2T3           10.   TM.1S2T3          -25.
    KIN1..T3  LY....T3            1.   KM.1..T3          -15.
    KIN1..T3  KN.1..T3            3.   FCC.....          68.1
    KIN1..T3  FIN...1.          300.   TYEAR1T3          -25.
    KIN1..T3  TYEAR2T3          -25.   LC.3..T3           -1.
    KIN1..T3  TYEAR3T3          -25.   LC.5..T3           -1.
    KVA1..T3  LM.1..T3          -10.   LN.1..T3            2.
    KVA1..T3  FCC.....        24.289   FIN...1.          107.
    KVA1..T3  LC.1..T3            1.
    KBU2S1T3  TN.2S1T3           10.   TM.2S1T3          -25.
    KBU2S2T3  TN.2S2T3           10.   TM.2S2T3          -25.
    KBU3S1T3  TN.3S1T3           10.   TM.3S1T3          -25.
    KBU3S2T3  TN.3S2T3           10.   TM.3S2T3          -25.
    KIN3..T3  LY....T3            1.   KM.3..T3          -15.
    KIN3..T3  KN.3..T3            3.   FCC.....           48.
    KIN3..T3  FIN...2.          300.   TYEAR2T3          -25.
    KIN3..T3  TYEAR3T3          -25.   LC.5..T3           -1.
    KVA3..T3  LM.3..T3          -10.   LN.3..T3            2.
    KVA3..T3  FCC.....         17.12   FIN...2.          107.
    KVA3..T3  LC.3..T3            1.
    KBU4S1T3  TN.4S1T3           10.   TM.4S1T3          -25.
    KBU4S2T3  TN.4S2T3           10.   TM.4S2T3          -25.
    KBU5S1T3  TN.5S1T3           10.   TM.5S1T3          -25.
    KBU5S2T3  TN.5S2T3           10.   TM.5S2T3          -25.
    KIN5..T3  LY....T3            1.   KM.5..T3          -15.
    KIN5..T3  KN.5..T3            3.   FCC.....          25.5
    KIN5..T3  FIN...3.          300.   TYEAR3T3          -25.
    KVA5..T3  LM.5..T3          -10.   LN.5..T3            2.
    KVA5..T3  FCC.....         9.095   FIN...3.          107.
    KVA5..T3  LC.5..T3            1.
    KBU6S1T3  TN.6S1T3           10.   TM.6S1T3          -25.
    KBU6S2T3  TN.6S2T3           10.   TM.6S2T3          -25.
    KBU1S1T4  TN.1S1T4           10.   TM.1S1T4          -25.
    KBU1S2T4  TN.1S2T4           10.   TM.1S2T4          -25.
    KIN1..T4  LY....T4            1.   KM.1..T4          -15.
    KIN1..T4  KN.1..T4            3.   FCC.....          68.1
    KIN1..T4  FIN...1.          300.   TYEAR1T4          -25.
    KIN1..T4  TYEAR2T4          -25.   LC.3..T4           -1.
    KIN1..T4  TYEAR3T4          -25.   LC.5..T4           -1.
    KVA1..T4  LM.1..T4          -10.   LN.1..T4            2.
    KVA1..T4  FCC.....        24.289   FIN...1.          107.
    KVA1..T4  LC.1..T4            1.
    KBU2S1T4  TN.2S1T4           10.   TM.2S1T4          -25.
    KBU2S2T4  TN.2S2T4           10.   TM.2S2T4          -25.
    KBU3S1T4  TN.3S1T4           10.   TM.3S1T4          -25.
    KBU3S2T4  TN.3S2T4           10.   TM.3S2T4          -25.
    KIN3..T4  LY....T4            1.   KM.3..T4          -15.
    KIN3..T4  KN.3..T4            3.   FCC.....           48.
    KIN3..T4  FIN...2.          300.   TYEAR2T4          -25.
    KIN3..T4  TYEAR3T4          -25.   LC.5..T4           -1.
    KVA3..T4  LM.3..T4          -10.   LN.3..T4            2.
    KVA3..T4  FCC.....         17.12   FIN...2.          107.
    KVA3..T4  LC.3..T4            1.
    KBU4S1T4  TN.4S1T4           10.   TM.4S1T4          -25.
    KBU4S2T4  TN.4S2T4           10.   TM.4S2T4          -25.
    KBU5S1T4  TN.5S1T4           10.   TM.5S1T4          -25.
    KBU5S2T4  TN.5S2T4           10.   TM.5S2T4          -25.
    KIN5..T4  LY....T4            1.   KM.5..T4          -15.
    KIN5..T4  KN.5..T4            3.   FCC.....          25.5
    KIN5..T4  FIN...3.          300.   TYEAR3T4          -25.
    KVA5..T4  LM.5..T4          -10.   LN.5..T4            2.
    KVA5..T4  FCC.....         9.095   FIN...3.          107.
    KVA5..T4  LC.5..T4            1.
    KBU6S1T4  TN.6S1T4           10.   TM.6S1T4          -25.
    KBU6S2T4  TN.6S2T4           10.   TM.6S2T4          -25.
    PEN...1.  FIN...1.           -1.   FAT...J.          100.
    PEN...2.  FIN...2.           -1.   FAT...J.          100.
    PEN...3.  FIN...3.           -1.   FAT...J.          100.
    RBU1S1T1  TB.1S1T1           -1.   TN.1S1T1           -1.
    RBU1S1T1  TM.1S1T1            1.   FTR.....         10.55
    RBU1S2T1  TB.1S2T1           -1.   TN.1S2T1           -1.
    RBU1S2T1  TM.1S2T1            1.   FTR.....           17.
    RPA1S1T1  TB.1S1T1           -1.   FTR.....         13.85
    RPA1S2T1  TB.1S2T1           -1.   FTR.....          17.5
    RBU2S1T1  TB.2S1T1           -1.   TN.2S1T1           -1.
    RBU2S1T1  TM.2S1T1            1.   FTR.....         10.55
    RBU2S2T1  TB.2S2T1           -1.   TN.2S2T1           -1.
    RBU2S2T1  TM.2S2T1            1.   FTR.....           17.
    RPA2S1T1  TB.2S1T1           -1.   FTR.....         13.85
    RPA2S2T1  TB.2S2T1           -1.   FTR.....          17.5
    RBU3S1T1  TB.3S1T1           -1.   TN.3S1T1           -1.
    RBU3S1T1  TM.3S1T1            1.   FTR.....         10.55
    RBU3S2T1  TB.3S2T1           -1.   TN.3S2T1           -1.
    RBU3S2T1  TM.3S2T1            1.   FTR.....           17.
    RPA3S1T1  TB.3S1T1           -1.   FTR.....         13.85
    RPA3S2T1  TB.3S2T1           -1.   FTR.....          17.5
    RBU4S1T1  TB.4S1T1           -1.   TN.4S1T1           -1.
    RBU4S1T1  TM.4S1T1            1.   FTR.....         10.55
    RBU4S2T1  TB.4S2T1           -1.   TN.4S2T1           -1.
    RBU4S2T1  TM.4S2T1            1.   FTR.....           17.
    RPA4S1T1  TB.4S1T1           -1.   FTR.....         13.85
    RPA4S2T1  TB.4S2T1           -1.   FTR.....          17.5
    RBU5S1T1  TB.5S1T1           -1.   TN.5S1T1           -1.
    RBU5S1T1  TM.5S1T1            1.   FTR.....         10.55
    RBU5S2T1  TB.5S2T1           -1.   TN.5S2T1           -1.
    RBU5S2T1  TM.5S2T1            1.   FTR.....           17.
    RPA5S1T1  TB.5S1T1           -1.   FTR.....         13.85
    RPA5S2T1  TB.5S2T1           -1.   FTR.....          17.5
    RBU6S1T1  TB.6S1T1           -1.   TN.6S1T1           -1.
    RBU6S1T1  TM.6S1T1            1.   FTR.....         10.55
    RBU6S2T1  TB.6S2T1           -1.   TN.6S2T1           -1.
    RBU6S2T1  TM.6S2T1            1.   FTR.....           17.
    RPA6S1T1  TB.6S1T1           -1.   FTR.....         13.85
    RPA6S2T1  TB.6S2T1           -1.   FTR.....          17.5
    RBU1S1T2  TB.1S1T2           -1.   TN.1S1T2           -1.
    RBU1S1T2  TM.1S1T2            1.   FTR.....         13.85
    RBU1S2T2  TB.1S2T2           -1.   TN.1S2T2           -1.
    RBU1S2T2  TM.1S2T2            1.   FTR.....         16.55
    RPA1S1T2  TB.1S1T2           -1.   FTR.....          16.3
    RPA1S2T2  TB.1S2T2           -1.   FTR.....           17.
    RBU2S1T2  TB.2S1T2           -1.   TN.2S1T2           -1.
    RBU2S1T2  TM.2S1T2            1.   FTR.....         13.85
    RBU2S2T2  TB.2S2T2           -1.   TN.2S2T2           -1.
    RBU2S2T2  TM.2S2T2            1.   FTR.....         16.55
    RPA2S1T2  TB.2S1T2           -1.   FTR.....          16.3
    RPA2S2T2  TB.2S2T2           -1.   FTR.....           17.
    RBU3S1T2  TB.3S1T2           -1.   TN.3S1T2           -1.
    RBU3S1T2  TM.3S1T2            1.   FTR.....         13.85
    RBU3S2T2  TB.3S2T2           -1.   TN.3S2T2           -1.
    RBU3S2T2  TM.3S2T2            1.   FTR.....         16.55
    RPA3S1T2  TB.3S1T2           -1.   FTR.....          16.3
    RPA3S2T2  TB.3S2T2           -1.   FTR.....           17.
    RBU4S1T2  TB.4S1T2           -1.   TN.4S1T2           -1.
    RBU4S1T2  TM.4S1T2            1.   FTR.....         13.85
    RBU4S2T2  TB.4S2T2           -1.   TN.4S2T2           -1.
    RBU4S2T2  TM.4S2T2            1.   FTR.....         16.55
    RPA4S1T2  TB.4S1T2           -1.   FTR.....          16.3
    RPA4S2T2  TB.4S2T2           -1.   FTR.....           17.
    RBU5S1T2  TB.5S1T2           -1.   TN.5S1T2           -1.
    RBU5S1T2  TM.5S1T2            1.   FTR.....         13.85
    RBU5S2T2  TB.5S2T2           -1.   TN.5S2T2           -1.
    RBU5S2T2  TM.5S2T2            1.   FTR.....         16.55
    RPA5S1T2  TB.5S1T2           -1.   FTR.....          16.3
    RPA5S2T2  TB.5S2T2           -1.   FTR.....           17.
    RBU6S1T2  TB.6S1T2           -1.   TN.6S1T2           -1.
    RBU6S1T2  TM.6S1T2            1.   FTR.....         13.85
    RBU6S2T2  TB.6S2T2           -1.   TN.6S2T2           -1.
    RBU6S2T2  TM.6S2T2            1.   FTR.....         16.55
    RPA6S1T2  TB.6S1T2           -1.   FTR.....          16.3
    RPA6S2T2  TB.6S2T2           -1.   FTR.....           17.
    RBU1S1T3  TB.1S1T3           -1.   TN.1S1T3           -1.
    RBU1S1T3  TM.1S1T3            1.   FTR.....         13.85
    RBU1S2T3  TB.1S2T3           -1.   TN.1S2T3           -1.
    RBU1S2T3  TM.1S2T3            1.   FTR.....         14.55
    RPA1S1T3  TB.1S1T3           -1.   FTR.....          16.3
    RPA1S2T3  TB.1S2T3           -1.   FTR.....           15.
    RBU2S1T3  TB.2S1T3           -1.   TN.2S1T3           -1.
    RBU2S1T3  TM.2S1T3            1.   FTR.....         13.85
    RBU2S2T3  TB.2S2T3           -1.   TN.2S2T3           -1.
    RBU2S2T3  TM.2S2T3            1.   FTR.....         14.55
    RPA2S1T3  TB.2S1T3           -1.   FTR.....          16.3
    RPA2S2T3  TB.2S2T3           -1.   FTR.....           15.
    RBU3S1T3  TB.3S1T3           -1.   TN.3S1T3           -1.
    RBU3S1T3  TM.3S1T3            1.   FTR.....         13.85
    RBU3S2T3  TB.3S2T3           -1.   TN.3S2T3           -1.
    RBU3S2T3  TM.3S2T3            1.   FTR.....         14.55
    RPA3S1T3  TB.3S1T3           -1.   FTR.....          16.3
    RPA3S2T3  TB.3S2T3           -1.   FTR.....           15.
    RBU4S1T3  TB.4S1T3           -1.   TN.4S1T3           -1.
    RBU4S1T3  TM.4S1T3            1.   FTR.....         13.85
    RBU4S2T3  TB.4S2T3           -1.   TN.4S2T3           -1.
    RBU4S2T3  TM.4S2T3            1.   FTR.....         14.55
    RPA4S1T3  TB.4S1T3           -1.   FTR.....          16.3
    RPA4S2T3  TB.4S2T3           -1.   FTR.....           15.
    RBU5S1T3  TB.5S1T3           -1.   TN.5S1T3           -1.
    RBU5S1T3  TM.5S1T3            1.   FTR.....         13.85
    RBU5S2T3  TB.5S2T3           -1.   TN.5S2T3           -1.
    RBU5S2T3  TM.5S2T3            1.   FTR.....         14.55
    RPA5S1T3  TB.5S1T3           -1.   FTR.....          16.3
    RPA5S2T3  TB.5S2T3           -1.   FTR.....           15.
    RBU6S1T3  TB.6S1T3           -1.   TN.6S1T3           -1.
    RBU6S1T3  TM.6S1T3            1.   FTR.....         13.85
    RBU6S2T3  TB.6S2T3           -1.   TN.6S2T3           -1.
    RBU6S2T3  TM.6S2T3            1.   FTR.....         14.55
    RPA6S1T3  TB.6S1T3           -1.   FTR.....          16.3
    RPA6S2T3  TB.6S2T3           -1.   FTR.....           15.
    RBU1S1T4  TB.1S1T4           -1.   TN.1S1T4           -1.
    RBU1S1T4  TM.1S1T4            1.   FTR.....         13.85
    RBU1S2T4  TB.1S2T4           -1.   TN.1S2T4           -1.
    RBU1S2T4  TM.1S2T4            1.   FTR.....         13.85
    RPA1S1T4  TB.1S1T4           -1.   FTR.....          16.3
    RPA1S2T4  TB.1S2T4           -1.   FTR.....         14.55
    RBU2S1T4  TB.2S1T4           -1.   TN.2S1T4           -1.
    RBU2S1T4  TM.2S1T4            1.   FTR.....         13.85
    RBU2S2T4  TB.2S2T4           -1.   TN.2S2T4           -1.
    RBU2S2T4  TM.2S2T4            1.   FTR.....         13.85
    RPA2S1T4  TB.2S1T4           -1.   FTR.....          16.3
    RPA2S2T4  TB.2S2T4           -1.   FTR.....         14.55
    RBU3S1T4  TB.3S1T4           -1.   TN.3S1T4           -1.
    RBU3S1T4  TM.3S1T4            1.   FTR.....         13.85
    RBU3S2T4  TB.3S2T4           -1.   TN.3S2T4           -1.
    RBU3S2T4  TM.3S2T4            1.   FTR.....         13.85
    RPA3S1T4  TB.3S1T4           -1.   FTR.....          16.3
    RPA3S2T4  TB.3S2T4           -1.   FTR.....         14.55
    RBU4S1T4  TB.4S1T4           -1.   TN.4S1T4           -1.
    RBU4S1T4  TM.4S1T4            1.   FTR.....         13.85
    RBU4S2T4  TB.4S2T4           -1.   TN.4S2T4           -1.
    RBU4S2T4  TM.4S2T4            1.   FTR.....         13.85
    RPA4S1T4  TB.4S1T4           -1.   FTR.....          16.3
    RPA4S2T4  TB.4S2T4           -1.   FTR.....         14.55
    RBU5S1T4  TB.5S1T4           -1.   TN.5S1T4           -1.
    RBU5S1T4  TM.5S1T4            1.   FTR.....         13.85
    RBU5S2T4  TB.5S2T4           -1.   TN.5S2T4           -1.
    RBU5S2T4  TM.5S2T4            1.   FTR.....         13.85
    RPA5S1T4  TB.5S1T4           -1.   FTR.....          16.3
    RPA5S2T4  TB.5S2T4           -1.   FTR.....         14.55
    RBU6S1T4  TB.6S1T4           -1.   TN.6S1T4           -1.
    RBU6S1T4  TM.6S1T4            1.   FTR.....         13.85
    RBU6S2T4  TB.6S2T4           -1.   TN.6S2T4           -1.
    RBU6S2T4  TM.6S2T4            1.   FTR.....         13.85
    RPA6S1T4  TB.6S1T4           -1.   FTR.....          16.3
    RPA6S2T4  TB.6S2T4           -1.   FTR.....         14.55
    ZP11S1T1  BP11..S1           -1.   BP11..T1            1.
    ZP11S1T1  TB.1S1T1            1.
    ZP11S2T1  BP11..S2           -1.   BP11..T1            1.
    ZP11S2T1  TB.1S2T1            1.
    ZP21S1T1  BP21..S1           -1.   BP21..T1            1.
    ZP21S1T1  TB.1S1T1            1.
    ZP21S2T1  BP21..S2           -1.   BP21..T1            1.
    ZP21S2T1  TB.1S2T1            1.
    ZP12S1T1  BP12..S1           -1.   BP12..T1            1.
    ZP12S1T1  TB.2S1T1            1.
    ZP12S2T1  BP12..S2           -1.   BP12..T1            1.
    ZP12S2T1  TB.2S2T1            1.
    ZP22S1T1  BP22..S1           -1.   BP22..T1            1.
    ZP22S1T1  TB.2S1T1            1.
    ZP22S2T1  BP22..S2           -1.   BP22..T1            1.
    ZP22S2T1  TB.2S2T1            1.
    ZP13S1T1  BP13..S1           -1.   BP13..T1            1.
    ZP13S1T1  TB.3S1T1            1.
    ZP13S2T1  BP13..S2           -1.   BP13..T1            1.
    ZP13S2T1  TB.3S2T1            1.
    ZP23S1T1  BP23..S1           -1.   BP23..T1            1.
    ZP23S1T1  TB.3S1T1            1.
    ZP23S2T1  BP23..S2           -1.   BP23..T1            1.
    ZP23S2T1  TB.3S2T1            1.
    ZP14S1T1  BP14..S1           -1.   BP14..T1            1.
    ZP14S1T1  TB.4S1T1            1.
    ZP14S2T1  BP14..S2           -1.   BP14..T1            1.
    ZP14S2T1  TB.4S2T1            1.
    ZP24S1T1  BP24..S1           -1.   BP24..T1            1.
    ZP24S1T1  TB.4S1T1            1.
    ZP24S2T1  BP24..S2           -1.   BP24..T1            1.
    ZP24S2T1  TB.4S2T1            1.
    ZP15S1T1  BP15..S1           -1.   BP15..T1            1.
    ZP15S1T1  TB.5S1T1            1.
    ZP15S2T1  BP15..S2           -1.   BP15..T1            1.
    ZP15S2T1  TB.5S2T1            1.
    ZP25S1T1  BP25..S1           -1.   BP25..T1            1.
    ZP25S1T1  TB.5S1T1            1.
    ZP25S2T1  BP25..S2           -1.   BP25..T1            1.
    ZP25S2T1  TB.5S2T1            1.
    ZP16S1T1  BP16..S1           -1.   BP16..T1            1.
    ZP16S1T1  TB.6S1T1            1.
    ZP16S2T1  BP16..S2           -1.   BP16..T1            1.
    ZP16S2T1  TB.6S2T1            1.
    ZP26S1T1  BP26..S1           -1.   BP26..T1            1.
    ZP26S1T1  TB.6S1T1            1.
    ZP26S2T1  BP26..S2           -1.   BP26..T1            1.
    ZP26S2T1  TB.6S2T1            1.
    ZP11S1T2  BP11..S1           -1.   BP11..T2            1.
    ZP11S1T2  TB.1S1T2            1.
    ZP11S2T2  BP11..S2           -1.   BP11..T2            1.
    ZP11S2T2  TB.1S2T2            1.
    ZP21S1T2  BP21..S1           -1.   BP21..T2            1.
    ZP21S1T2  TB.1S1T2            1.
    ZP21S2T2  BP21..S2           -1.   BP21..T2            1.
    ZP21S2T2  TB.1S2T2            1.
    ZP12S1T2  BP12..S1           -1.   BP12..T2            1.
    ZP12S1T2  TB.2S1T2            1.
    ZP12S2T2  BP12..S2           -1.   BP12..T2            1.
    ZP12S2T2  TB.2S2T2            1.
    ZP22S1T2  BP22..S1           -1.   BP22..T2            1.
    ZP22S1T2  TB.2S1T2            1.
    ZP22S2T2  BP22..S2           -1.   BP22..T2            1.
    ZP22S2T2  TB.2S2T2            1.
    ZP13S1T2  BP13..S1           -1.   BP13..T2            1.
    ZP13S1T2  TB.3S1T2            1.
    ZP13S2T2  BP13..S2           -1.   BP13..T2            1.
    ZP13S2T2  TB.3S2T2            1.
    ZP23S1T2  BP23..S1           -1.   BP23..T2            1.
    ZP23S1T2  TB.3S1T2            1.
    ZP23S2T2  BP23..S2           -1.   BP23..T2            1.
    ZP23S2T2  TB.3S2T2            1.
    ZP14S1T2  BP14..S1           -1.   BP14..T2            1.
    ZP14S1T2  TB.4S1T2            1.
    ZP14S2T2  BP14..S2           -1.   BP14..T2            1.
    ZP14S2T2  TB.4S2T2            1.
    ZP24S1T2  BP24..S1           -1.   BP24..T2            1.
    ZP24S1T2  TB.4S1T2            1.
    ZP24S2T2  BP24..S2           -1.   BP24..T2            1.
    ZP24S2T2  TB.4S2T2            1.
    ZP15S1T2  BP15..S1           -1.   BP15..T2            1.
    ZP15S1T2  TB.5S1T2            1.
    ZP15S2T2  BP15..S2           -1.   BP15..T2            1.
    ZP15S2T2  TB.5S2T2            1.
    ZP25S1T2  BP25..S1           -1.   BP25..T2            1.
    ZP25S1T2  TB.5S1T2            1.
    ZP25S2T2  BP25..S2           -1.   BP25..T2            1.
    ZP25S2T2  TB.5S2T2            1.
    ZP16S1T2  BP16..S1           -1.   BP16..T2            1.
    ZP16S1T2  TB.6S1T2            1.
    ZP16S2T2  BP16..S2           -1.   BP16..T2            1.
    ZP16S2T2  TB.6S2T2            1.
    ZP26S1T2  BP26..S1           -1.   BP26..T2            1.
    ZP26S1T2  TB.6S1T2            1.
    ZP26S2T2  BP26..S2           -1.   BP26..T2            1.
    ZP26S2T2  TB.6S2T2            1.
    ZP11S1T3  BP11..S1           -1.   BP11..T3            1.
    ZP11S1T3  TB.1S1T3            1.
    ZP11S2T3  BP11..S2           -1.   BP11..T3            1.
    ZP11S2T3  TB.1S2T3            1.
    ZP21S1T3  BP21..S1           -1.   BP21..T3            1.
    ZP21S1T3  TB.1S1T3            1.
    ZP21S2T3  BP21..S2           -1.   BP21..T3            1.
    ZP21S2T3  TB.1S2T3            1.
    ZP12S1T3  BP12..S1           -1.   BP12..T3            1.
    ZP12S1T3  TB.2S1T3            1.
    ZP12S2T3  BP12..S2           -1.   BP12..T3            1.
    ZP12S2T3  TB.2S2T3            1.
    ZP22S1T3  BP22..S1           -1.   BP22..T3            1.
    ZP22S1T3  TB.2S1T3            1.
    ZP22S2T3  BP22..S2           -1.   BP22..T3            1.
    ZP22S2T3  TB.2S2T3            1.
    ZP13S1T3  BP13..S1           -1.   BP13..T3            1.
    ZP13S1T3  TB.3S1T3            1.
    ZP13S2T3  BP13..S2           -1.   BP13..T3            1.
    ZP13S2T3  TB.3S2T3            1.
    ZP23S1T3  BP23..S1           -1.   BP23..T3            1.
    ZP23S1T3  TB.3S1T3            1.
    ZP23S2T3  BP23..S2           -1.   BP23..T3            1.
    ZP23S2T3  TB.3S2T3            1.
    ZP14S1T3  BP14..S1           -1.   BP14..T3            1.
    ZP14S1T3  TB.4S1T3            1.
    ZP14S2T3  BP14..S2           -1.   BP14..T3            1.
    ZP14S2T3  TB.4S2T3            1.
    ZP24S1T3  BP24..S1           -1.   BP24..T3            1.
    ZP24S1T3  TB.4S1T3            1.
    ZP24S2T3  BP24..S2           -1.   BP24..T3            1.
    ZP24S2T3  TB.4S2T3            1.
    ZP15S1T3  BP15..S1           -1.   BP15..T3            1.
    ZP15S1T3  TB.5S1T3            1.
    ZP15S2T3  BP15..S2           -1.   BP15..T3            1.
    ZP15S2T3  TB.5S2T3            1.
    ZP25S1T3  BP25..S1           -1.   BP25..T3            1.
    ZP25S1T3  TB.5S1T3            1.
    ZP25S2T3  BP25..S2           -1.   BP25..T3            1.
    ZP25S2T3  TB.5S2T3            1.
    ZP16S1T3  BP16..S1           -1.   BP16..T3            1.
    ZP16S1T3  TB.6S1T3            1.
    ZP16S2T3  BP16..S2           -1.   BP16..T3            1.
    ZP16S2T3  TB.6S2T3            1.
    ZP26S1T3  BP26..S1           -1.   BP26..T3            1.
    ZP26S1T3  TB.6S1T3            1.
    ZP26S2T3  BP26..S2           -1.   BP26..T3            1.
    ZP26S2T3  TB.6S2T3            1.
    ZP11S1T4  BP11..S1           -1.   BP11..T4            1.
    ZP11S1T4  TB.1S1T4            1.
    ZP11S2T4  BP11..S2           -1.   BP11..T4            1.
    ZP11S2T4  TB.1S2T4            1.
    ZP21S1T4  BP21..S1           -1.   BP21..T4            1.
    ZP21S1T4  TB.1S1T4            1.
    ZP21S2T4  BP21..S2           -1.   BP21..T4            1.
    ZP21S2T4  TB.1S2T4            1.
    ZP12S1T4  BP12..S1           -1.   BP12..T4            1.
    ZP12S1T4  TB.2S1T4            1.
    ZP12S2T4  BP12..S2           -1.   BP12..T4            1.
    ZP12S2T4  TB.2S2T4            1.
    ZP22S1T4  BP22..S1           -1.   BP22..T4            1.
    ZP22S1T4  TB.2S1T4            1.
    ZP22S2T4  BP22..S2           -1.   BP22..T4            1.
    ZP22S2T4  TB.2S2T4            1.
    ZP13S1T4  BP13..S1           -1.   BP13..T4            1.
    ZP13S1T4  TB.3S1T4            1.
    ZP13S2T4  BP13..S2           -1.   BP13..T4            1.
    ZP13S2T4  TB.3S2T4            1.
    ZP23S1T4  BP23..S1           -1.   BP23..T4            1.
    ZP23S1T4  TB.3S1T4            1.
    ZP23S2T4  BP23..S2           -1.   BP23..T4            1.
    ZP23S2T4  TB.3S2T4            1.
    ZP14S1T4  BP14..S1           -1.   BP14..T4            1.
    ZP14S1T4  TB.4S1T4            1.
    ZP14S2T4  BP14..S2           -1.   BP14..T4            1.
    ZP14S2T4  TB.4S2T4            1.
    ZP24S1T4  BP24..S1           -1.   BP24..T4            1.
    ZP24S1T4  TB.4S1T4            1.
    ZP24S2T4  BP24..S2           -1.   BP24..T4            1.
    ZP24S2T4  TB.4S2T4            1.
    ZP15S1T4  BP15..S1           -1.   BP15..T4            1.
    ZP15S1T4  TB.5S1T4            1.
    ZP15S2T4  BP15..S2           -1.   BP15..T4            1.
    ZP15S2T4  TB.5S2T4            1.
    ZP25S1T4  BP25..S1           -1.   BP25..T4            1.
    ZP25S1T4  TB.5S1T4            1.
    ZP25S2T4  BP25..S2           -1.   BP25..T4            1.
    ZP25S2T4  TB.5S2T4            1.
    ZP16S1T4  BP16..S1           -1.   BP16..T4            1.
    ZP16S1T4  TB.6S1T4            1.
    ZP16S2T4  BP16..S2           -1.   BP16..T4            1.
    ZP16S2T4  TB.6S2T4            1.
    ZP26S1T4  BP26..S1           -1.   BP26..T4            1.
    ZP26S1T4  TB.6S1T4            1.
    ZP26S2T4  BP26..S2           -1.   BP26..T4            1.
    ZP26S2T4  TB.6S2T4            1.
    ZP11S101  BP11..S1           -1.   FTR.....          28.2
    ZP11S101  ZP11..01            1.
    ZP11S201  BP11..S2           -1.   FTR.....          34.3
    ZP11S201  ZP11..01            1.
    ZP11T101  BP11..T1           -1.   FTR.....           8.2
    ZP11T101  TYEAR1T1            1.   ZP11..01            1.
    ZP11T201  BP11..T2           -1.   FTR.....         13.25
    ZP11T201  TYEAR1T2            1.   ZP11..01            1.
    ZP11T301  BP11..T3           -1.   FTR.....          15.8
    ZP11T301  TYEAR1T3            1.   ZP11..01            1.
    ZP11T401  BP11..T4           -1.   FTR.....         17.95
    ZP11T401  TYEAR1T4            1.   ZP11..01            1.
    ZP21S101  BP21..S1           -1.   FTR.....          28.2
    ZP21S101  ZP21..01            1.
    ZP21S201  BP21..S2           -1.   FTR.....          34.3
    ZP21S201  ZP21..01            1.
    ZP21T101  BP21..T1           -1.   FTR.....           8.2
    ZP21T101  TYEAR1T1            1.   ZP21..01            1.
    ZP21T201  BP21..T2           -1.   FTR.....         13.25
    ZP21T201  TYEAR1T2            1.   ZP21..01            1.
    ZP21T301  BP21..T3           -1.   FTR.....          15.8
    ZP21T301  TYEAR1T3            1.   ZP21..01            1.
    ZP21T401  BP21..T4           -1.   FTR.....         17.95
    ZP21T401  TYEAR1T4            1.   ZP21..01            1.
    ZP12S101  BP12..S1           -1.   FTR.....          28.2
    ZP12S101  ZP12..01            1.
    ZP12S201  BP12..S2           -1.   FTR.....          34.3
    ZP12S201  ZP12..01            1.
    ZP12T101  BP12..T1           -1.   FTR.....           8.2
    ZP12T101  TYEAR1T1            1.   ZP12..01            1.
    ZP12T201  BP12..T2           -1.   FTR.....         13.25
    ZP12T201  TYEAR1T2            1.   ZP12..01            1.
    ZP12T301  BP12..T3           -1.   FTR.....          15.8
    ZP12T301  TYEAR1T3            1.   ZP12..01            1.
    ZP12T401  BP12..T4           -1.   FTR.....         17.95
    ZP12T401  TYEAR1T4            1.   ZP12..01            1.
    ZP22S101  BP22..S1           -1.   FTR.....          28.2
    ZP22S101  ZP22..01            1.
    ZP22S201  BP22..S2           -1.   FTR.....          34.3
    ZP22S201  ZP22..01            1.
    ZP22T101  BP22..T1           -1.   FTR.....           8.2
    ZP22T101  TYEAR1T1            1.   ZP22..01            1.
    ZP22T201  BP22..T2           -1.   FTR.....         13.25
    ZP22T201  TYEAR1T2            1.   ZP22..01            1.
    ZP22T301  BP22..T3           -1.   FTR.....          15.8
    ZP22T301  TYEAR1T3            1.   ZP22..01            1.
    ZP22T401  BP22..T4           -1.   FTR.....         17.95
    ZP22T401  TYEAR1T4            1.   ZP22..01            1.
    ZP13S101  BP13..S1           -1.   FTR.....          28.2
    ZP13S101  ZP13..01            1.
    ZP13S201  BP13..S2           -1.   FTR.....          34.3
    ZP13S201  ZP13..01            1.
    ZP13T101  BP13..T1           -1.   FTR.....           8.2
    ZP13T101  TYEAR2T1            1.   ZP13..01            1.
    ZP13T201  BP13..T2           -1.   FTR.....         13.25
    ZP13T201  TYEAR2T2            1.   ZP13..01            1.
    ZP13T301  BP13..T3           -1.   FTR.....          15.8
    ZP13T301  TYEAR2T3            1.   ZP13..01            1.
    ZP13T401  BP13..T4           -1.   FTR.....         17.95
    ZP13T401  TYEAR2T4            1.   ZP13..01            1.
    ZP23S101  BP23..S1           -1.   FTR.....          28.2
    ZP23S101  ZP23..01            1.
    ZP23S201  BP23..S2           -1.   FTR.....          34.3
    ZP23S201  ZP23..01            1.
    ZP23T101  BP23..T1           -1.   FTR.....           8.2
    ZP23T101  TYEAR2T1            1.   ZP23..01            1.
    ZP23T201  BP23..T2           -1.   FTR.....         13.25
    ZP23T201  TYEAR2T2            1.   ZP23..01            1.
    ZP23T301  BP23..T3           -1.   FTR.....          15.8
    ZP23T301  TYEAR2T3            1.   ZP23..01            1.
    ZP23T401  BP23..T4           -1.   FTR.....         17.95
    ZP23T401  TYEAR2T4            1.   ZP23..01            1.
    ZP14S101  BP14..S1           -1.   FTR.....          28.2
    ZP14S101  ZP14..01            1.
    ZP14S201  BP14..S2           -1.   FTR.....          34.3
    ZP14S201  ZP14..01            1.
    ZP14T101  BP14..T1           -1.   FTR.....           8.2
    ZP14T101  TYEAR2T1            1.   ZP14..01            1.
    ZP14T201  BP14..T2           -1.   FTR.....         13.25
    ZP14T201  TYEAR2T2            1.   ZP14..01            1.
    ZP14T301  BP14..T3           -1.   FTR.....          15.8
    ZP14T301  TYEAR2T3            1.   ZP14..01            1.
    ZP14T401  BP14..T4           -1.   FTR.....         17.95
    ZP14T401  TYEAR2T4            1.   ZP14..01            1.
    ZP24S101  BP24..S1           -1.   FTR.....          28.2
    ZP24S101  ZP24..01            1.
    ZP24S201  BP24..S2           -1.   FTR.....          34.3
    ZP24S201  ZP24..01            1.
    ZP24T101  BP24..T1           -1.   FTR.....           8.2
    ZP24T101  TYEAR2T1            1.   ZP24..01            1.
    ZP24T201  BP24..T2           -1.   FTR.....         13.25
    ZP24T201  TYEAR2T2            1.   ZP24..01            1.
    ZP24T301  BP24..T3           -1.   FTR.....          15.8
    ZP24T301  TYEAR2T3            1.   ZP24..01            1.
    ZP24T401  BP24..T4           -1.   FTR.....         17.95
    ZP24T401  TYEAR2T4            1.   ZP24..01            1.
    ZP15S101  BP15..S1           -1.   FTR.....          28.2
    ZP15S101  ZP15..01            1.
    ZP15S201  BP15..S2           -1.   FTR.....          34.3
    ZP15S201  ZP15..01            1.
    ZP15T101  BP15..T1           -1.   FTR.....           8.2
    ZP15T101  TYEAR3T1            1.   ZP15..01            1.
    ZP15T201  BP15..T2           -1.   FTR.....         13.25
    ZP15T201  TYEAR3T2            1.   ZP15..01            1.
    ZP15T301  BP15..T3           -1.   FTR.....          15.8
    ZP15T301  TYEAR3T3            1.   ZP15..01            1.
    ZP15T401  BP15..T4           -1.   FTR.....         17.95
    ZP15T401  TYEAR3T4            1.   ZP15..01            1.
    ZP25S101  BP25..S1           -1.   FTR.....          28.2
    ZP25S101  ZP25..01            1.
    ZP25S201  BP25..S2           -1.   FTR.....          34.3
    ZP25S201  ZP25..01            1.
    ZP25T101  BP25..T1           -1.   FTR.....           8.2
    ZP25T101  TYEAR3T1            1.   ZP25..01            1.
    ZP25T201  BP25..T2           -1.   FTR.....         13.25
    ZP25T201  TYEAR3T2            1.   ZP25..01            1.
    ZP25T301  BP25..T3           -1.   FTR.....          15.8
    ZP25T301  TYEAR3T3            1.   ZP25..01            1.
    ZP25T401  BP25..T4           -1.   FTR.....         17.95
    ZP25T401  TYEAR3T4            1.   ZP25..01            1.
    ZP16S101  BP16..S1           -1.   FTR.....          28.2
    ZP16S101  ZP16..01            1.
    ZP16S201  BP16..S2           -1.   FTR.....          34.3
    ZP16S201  ZP16..01            1.
    ZP16T101  BP16..T1           -1.   FTR.....           8.2
    ZP16T101  TYEAR3T1            1.   ZP16..01            1.
    ZP16T201  BP16..T2           -1.   FTR.....         13.25
    ZP16T201  TYEAR3T2            1.   ZP16..01            1.
    ZP16T301  BP16..T3           -1.   FTR.....          15.8
    ZP16T301  TYEAR3T3            1.   ZP16..01            1.
    ZP16T401  BP16..T4           -1.   FTR.....         17.95
    ZP16T401  TYEAR3T4            1.   ZP16..01            1.
    ZP26S101  BP26..S1           -1.   FTR.....          28.2
    ZP26S101  ZP26..01            1.
    ZP26S201  BP26..S2           -1.   FTR.....          34.3
    ZP26S201  ZP26..01            1.
    ZP26T101  BP26..T1           -1.   FTR.....           8.2
    ZP26T101  TYEAR3T1            1.   ZP26..01            1.
    ZP26T201  BP26..T2           -1.   FTR.....         13.25
    ZP26T201  TYEAR3T2            1.   ZP26..01            1.
    ZP26T301  BP26..T3           -1.   FTR.....          15.8
    ZP26T301  TYEAR3T3            1.   ZP26..01            1.
    ZP26T401  BP26..T4           -1.   FTR.....         17.95
    ZP26T401  TYEAR3T4            1.   ZP26..01            1.
    ZP11S102  BP11..S1           -1.   FTR.....          24.4
    ZP11S102  ZP11..02            1.
    ZP11S202  BP11..S2           -1.   FTR.....          33.4
    ZP11S202  ZP11..02            1.
    ZP11T102  BP11..T1           -1.   FTR.....         11.85
    ZP11T102  TYEAR1T1            1.   ZP11..02            1.
    ZP11T202  BP11..T2           -1.   FTR.....         14.55
    ZP11T202  TYEAR1T2            1.   ZP11..02            1.
    ZP11T302  BP11..T3           -1.   FTR.....          13.9
    ZP11T302  TYEAR1T3            1.   ZP11..02            1.
    ZP11T402  BP11..T4           -1.   FTR.....         17.15
    ZP11T402  TYEAR1T4            1.   ZP11..02            1.
    ZP21S102  BP21..S1           -1.   FTR.....          24.4
    ZP21S102  ZP21..02            1.
    ZP21S202  BP21..S2           -1.   FTR.....          33.4
    ZP21S202  ZP21..02            1.
    ZP21T102  BP21..T1           -1.   FTR.....         11.85
    ZP21T102  TYEAR1T1            1.   ZP21..02            1.
    ZP21T202  BP21..T2           -1.   FTR.....         14.55
    ZP21T202  TYEAR1T2            1.   ZP21..02            1.
    ZP21T302  BP21..T3           -1.   FTR.....          13.9
    ZP21T302  TYEAR1T3            1.   ZP21..02            1.
    ZP21T402  BP21..T4           -1.   FTR.....         17.15
    ZP21T402  TYEAR1T4            1.   ZP21..02            1.
    ZP12S102  BP12..S1           -1.   FTR.....          24.4
    ZP12S102  ZP12..02            1.
    ZP12S202  BP12..S2           -1.   FTR.....          33.4
    ZP12S202  ZP12..02            1.
    ZP12T102  BP12..T1           -1.   FTR.....         11.85
    ZP12T102  TYEAR1T1            1.   ZP12..02            1.
    ZP12T202  BP12..T2           -1.   FTR.....         14.55
    ZP12T202  TYEAR1T2            1.   ZP12..02            1.
    ZP12T302  BP12..T3           -1.   FTR.....          13.9
    ZP12T302  TYEAR1T3            1.   ZP12..02            1.
    ZP12T402  BP12..T4           -1.   FTR.....         17.15
    ZP12T402  TYEAR1T4            1.   ZP12..02            1.
    ZP22S102  BP22..S1           -1.   FTR.....          24.4
    ZP22S102  ZP22..02            1.
    ZP22S202  BP22..S2           -1.   FTR.....          33.4
    ZP22S202  ZP22..02            1.
    ZP22T102  BP22..T1           -1.   FTR.....         11.85
    ZP22T102  TYEAR1T1            1.   ZP22..02            1.
    ZP22T202  BP22..T2           -1.   FTR.....         14.55
    ZP22T202  TYEAR1T2            1.   ZP22..02            1.
    ZP22T302  BP22..T3           -1.   FTR.....          13.9
    ZP22T302  TYEAR1T3            1.   ZP22..02            1.
    ZP22T402  BP22..T4           -1.   FTR.....         17.15
    ZP22T402  TYEAR1T4            1.   ZP22..02            1.
    ZP13S102  BP13..S1           -1.   FTR.....          24.4
    ZP13S102  ZP13..02            1.
    ZP13S202  BP13..S2           -1.   FTR.....          33.4
    ZP13S202  ZP13..02            1.
    ZP13T102  BP13..T1           -1.   FTR.....         11.85
    ZP13T102  TYEAR2T1            1.   ZP13..02            1.
    ZP13T202  BP13..T2           -1.   FTR.....         14.55
    ZP13T202  TYEAR2T2            1.   ZP13..02            1.
    ZP13T302  BP13..T3           -1.   FTR.....          13.9
    ZP13T302  TYEAR2T3            1.   ZP13..02            1.
    ZP13T402  BP13..T4           -1.   FTR.....         17.15
    ZP13T402  TYEAR2T4            1.   ZP13..02            1.
    ZP23S102  BP23..S1           -1.   FTR.....          24.4
    ZP23S102  ZP23..02            1.
    ZP23S202  BP23..S2           -1.   FTR.....          33.4
    ZP23S202  ZP23..02            1.
    ZP23T102  BP23..T1           -1.   FTR.....         11.85
    ZP23T102  TYEAR2T1            1.   ZP23..02            1.
    ZP23T202  BP23..T2           -1.   FTR.....         14.55
    ZP23T202  TYEAR2T2            1.   ZP23..02            1.
    ZP23T302  BP23..T3           -1.   FTR.....          13.9
    ZP23T302  TYEAR2T3            1.   ZP23..02            1.
    ZP23T402  BP23..T4           -1.   FTR.....         17.15
    ZP23T402  TYEAR2T4            1.   ZP23..02            1.
    ZP14S102  BP14..S1           -1.   FTR.....          24.4
    ZP14S102  ZP14..02            1.
    ZP14S202  BP14..S2           -1.   FTR.....          33.4
    ZP14S202  ZP14..02            1.
    ZP14T102  BP14..T1           -1.   FTR.....         11.85
    ZP14T102  TYEAR2T1            1.   ZP14..02            1.
    ZP14T202  BP14..T2           -1.   FTR.....         14.55
    ZP14T202  TYEAR2T2            1.   ZP14..02            1.
    ZP14T302  BP14..T3           -1.   FTR.....          13.9
    ZP14T302  TYEAR2T3            1.   ZP14..02            1.
    ZP14T402  BP14..T4           -1.   FTR.....         17.15
    ZP14T402  TYEAR2T4            1.   ZP14..02            1.
    ZP24S102  BP24..S1           -1.   FTR.....          24.4
    ZP24S102  ZP24..02            1.
    ZP24S202  BP24..S2           -1.   FTR.....          33.4
    ZP24S202  ZP24..02            1.
    ZP24T102  BP24..T1           -1.   FTR.....         11.85
    ZP24T102  TYEAR2T1            1.   ZP24..02            1.
    ZP24T202  BP24..T2           -1.   FTR.....         14.55
    ZP24T202  TYEAR2T2            1.   ZP24..02            1.
    ZP24T302  BP24..T3           -1.   FTR.....          13.9
    ZP24T302  TYEAR2T3            1.   ZP24..02            1.
    ZP24T402  BP24..T4           -1.   FTR.....         17.15
    ZP24T402  TYEAR2T4            1.   ZP24..02            1.
    ZP15S102  BP15..S1           -1.   FTR.....          24.4
    ZP15S102  ZP15..02            1.
    ZP15S202  BP15..S2           -1.   FTR.....          33.4
    ZP15S202  ZP15..02            1.
    ZP15T102  BP15..T1           -1.   FTR.....         11.85
    ZP15T102  TYEAR3T1            1.   ZP15..02            1.
    ZP15T202  BP15..T2           -1.   FTR.....         14.55
    ZP15T202  TYEAR3T2            1.   ZP15..02            1.
    ZP15T302  BP15..T3           -1.   FTR.....          13.9
    ZP15T302  TYEAR3T3            1.   ZP15..02            1.
    ZP15T402  BP15..T4           -1.   FTR.....         17.15
    ZP15T402  TYEAR3T4            1.   ZP15..02            1.
    ZP25S102  BP25..S1           -1.   FTR.....          24.4
    ZP25S102  ZP25..02            1.
    ZP25S202  BP25..S2           -1.   FTR.....          33.4
    ZP25S202  ZP25..02            1.
    ZP25T102  BP25..T1           -1.   FTR.....         11.85
    ZP25T102  TYEAR3T1            1.   ZP25..02            1.
    ZP25T202  BP25..T2           -1.   FTR.....         14.55
    ZP25T202  TYEAR3T2            1.   ZP25..02            1.
    ZP25T302  BP25..T3           -1.   FTR.....          13.9
    ZP25T302  TYEAR3T3            1.   ZP25..02            1.
    ZP25T402  BP25..T4           -1.   FTR.....         17.15
    ZP25T402  TYEAR3T4            1.   ZP25..02            1.
    ZP16S102  BP16..S1           -1.   FTR.....          24.4
    ZP16S102  ZP16..02            1.
    ZP16S202  BP16..S2           -1.   FTR.....          33.4
    ZP16S202  ZP16..02            1.
    ZP16T102  BP16..T1           -1.   FTR.....         11.85
    ZP16T102  TYEAR3T1            1.   ZP16..02            1.
    ZP16T202  BP16..T2           -1.   FTR.....         14.55
    ZP16T202  TYEAR3T2            1.   ZP16..02            1.
    ZP16T302  BP16..T3           -1.   FTR.....          13.9
    ZP16T302  TYEAR3T3            1.   ZP16..02            1.
    ZP16T402  BP16..T4           -1.   FTR.....         17.15
    ZP16T402  TYEAR3T4            1.   ZP16..02            1.
    ZP26S102  BP26..S1           -1.   FTR.....          24.4
    ZP26S102  ZP26..02            1.
    ZP26S202  BP26..S2           -1.   FTR.....          33.4
    ZP26S202  ZP26..02            1.
    ZP26T102  BP26..T1           -1.   FTR.....         11.85
    ZP26T102  TYEAR3T1            1.   ZP26..02            1.
    ZP26T202  BP26..T2           -1.   FTR.....         14.55
    ZP26T202  TYEAR3T2            1.   ZP26..02            1.
    ZP26T302  BP26..T3           -1.   FTR.....          13.9
    ZP26T302  TYEAR3T3            1.   ZP26..02            1.
    ZP26T402  BP26..T4           -1.   FTR.....         17.15
    ZP26T402  TYEAR3T4            1.   ZP26..02            1.
    ZP11S103  BP11..S1           -1.   FTR.....          25.4
    ZP11S103  ZP11..03            1.
    ZP11S203  BP11..S2           -1.   FTR.....         26.25
    ZP11S203  ZP11..03            1.
    ZP11T103  BP11..T1           -1.   FTR.....           7.6
    ZP11T103  TYEAR1T1            1.   ZP11..03            1.
    ZP11T203  BP11..T2           -1.   FTR.....           8.7
    ZP11T203  TYEAR1T2            1.   ZP11..03            1.
    ZP11T303  BP11..T3           -1.   FTR.....          8.35
    ZP11T303  TYEAR1T3            1.   ZP11..03            1.
    ZP11T403  BP11..T4           -1.   FTR.....         10.45
    ZP11T403  TYEAR1T4            1.   ZP11..03            1.
    ZP21S103  BP21..S1           -1.   FTR.....          25.4
    ZP21S103  ZP21..03            1.
    ZP21S203  BP21..S2           -1.   FTR.....         26.25
    ZP21S203  ZP21..03            1.
    ZP21T103  BP21..T1           -1.   FTR.....           7.6
    ZP21T103  TYEAR1T1            1.   ZP21..03            1.
    ZP21T203  BP21..T2           -1.   FTR.....           8.7
    ZP21T203  TYEAR1T2            1.   ZP21..03            1.
    ZP21T303  BP21..T3           -1.   FTR.....          8.35
    ZP21T303  TYEAR1T3            1.   ZP21..03            1.
    ZP21T403  BP21..T4           -1.   FTR.....         10.45
    ZP21T403  TYEAR1T4            1.   ZP21..03            1.
    ZP12S103  BP12..S1           -1.   FTR.....          25.4
    ZP12S103  ZP12..03            1.
    ZP12S203  BP12..S2           -1.   FTR.....         26.25
    ZP12S203  ZP12..03            1.
    ZP12T103  BP12..T1           -1.   FTR.....           7.6
    ZP12T103  TYEAR1T1            1.   ZP12..03            1.
    ZP12T203  BP12..T2           -1.   FTR.....           8.7
    ZP12T203  TYEAR1T2            1.   ZP12..03            1.
    ZP12T303  BP12..T3           -1.   FTR.....          8.35
    ZP12T303  TYEAR1T3            1.   ZP12..03            1.
    ZP12T403  BP12..T4           -1.   FTR.....         10.45
    ZP12T403  TYEAR1T4            1.   ZP12..03            1.
    ZP22S103  BP22..S1           -1.   FTR.....          25.4
    ZP22S103  ZP22..03            1.
    ZP22S203  BP22..S2           -1.   FTR.....         26.25
    ZP22S203  ZP22..03            1.
    ZP22T103  BP22..T1           -1.   FTR.....           7.6
    ZP22T103  TYEAR1T1            1.   ZP22..03            1.
    ZP22T203  BP22..T2           -1.   FTR.....           8.7
    ZP22T203  TYEAR1T2            1.   ZP22..03            1.
    ZP22T303  BP22..T3           -1.   FTR.....          8.35
    ZP22T303  TYEAR1T3            1.   ZP22..03            1.
    ZP22T403  BP22..T4           -1.   FTR.....         10.45
    ZP22T403  TYEAR1T4            1.   ZP22..03            1.
    ZP13S103  BP13..S1           -1.   FTR.....          25.4
    ZP13S103  ZP13..03            1.
    ZP13S203  BP13..S2           -1.   FTR.....         26.25
    ZP13S203  ZP13..03            1.
    ZP13T103  BP13..T1           -1.   FTR.....           7.6
    ZP13T103  TYEAR2T1            1.   ZP13..03            1.
    ZP13T203  BP13..T2           -1.   FTR.....           8.7
    ZP13T203  TYEAR2T2            1.   ZP13..03            1.
    ZP13T303  BP13..T3           -1.   FTR.....          8.35
    ZP13T303  TYEAR2T3            1.   ZP13..03            1.
    ZP13T403  BP13..T4           -1.   FTR.....         10.45
    ZP13T403  TYEAR2T4            1.   ZP13..03            1.
    ZP23S103  BP23..S1           -1.   FTR.....          25.4
    ZP23S103  ZP23..03            1.
    ZP23S203  BP23..S2           -1.   FTR.....         26.25
    ZP23S203  ZP23..03            1.
    ZP23T103  BP23..T1           -1.   FTR.....           7.6
    ZP23T103  TYEAR2T1            1.   ZP23..03            1.
    ZP23T203  BP23..T2           -1.   FTR.....           8.7
    ZP23T203  TYEAR2T2            1.   ZP23..03            1.
    ZP23T303  BP23..T3           -1.   FTR.....          8.35
    ZP23T303  TYEAR2T3            1.   ZP23..03            1.
    ZP23T403  BP23..T4           -1.   FTR.....         10.45
    ZP23T403  TYEAR2T4            1.   ZP23..03            1.
    ZP14S103  BP14..S1           -1.   FTR.....          25.4
    ZP14S103  ZP14..03            1.
    ZP14S203  BP14..S2           -1.   FTR.....         26.25
    ZP14S203  ZP14..03            1.
    ZP14T103  BP14..T1           -1.   FTR.....           7.6
    ZP14T103  TYEAR2T1            1.   ZP14..03            1.
    ZP14T203  BP14..T2           -1.   FTR.....           8.7
    ZP14T203  TYEAR2T2            1.   ZP14..03            1.
    ZP14T303  BP14..T3           -1.   FTR.....          8.35
    ZP14T303  TYEAR2T3            1.   ZP14..03            1.
    ZP14T403  BP14..T4           -1.   FTR.....         10.45
    ZP14T403  TYEAR2T4            1.   ZP14..03            1.
    ZP24S103  BP24..S1           -1.   FTR.....          25.4
    ZP24S103  ZP24..03            1.
    ZP24S203  BP24..S2           -1.   FTR.....         26.25
    ZP24S203  ZP24..03            1.
    ZP24T103  BP24..T1           -1.   FTR.....           7.6
    ZP24T103  TYEAR2T1            1.   ZP24..03            1.
    ZP24T203  BP24..T2           -1.   FTR.....           8.7
    ZP24T203  TYEAR2T2            1.   ZP24..03            1.
    ZP24T303  BP24..T3           -1.   FTR.....          8.35
    ZP24T303  TYEAR2T3            1.   ZP24..03            1.
    ZP24T403  BP24..T4           -1.   FTR.....         10.45
    ZP24T403  TYEAR2T4            1.   ZP24..03            1.
    ZP15S103  BP15..S1           -1.   FTR.....          25.4
    ZP15S103  ZP15..03            1.
    ZP15S203  BP15..S2           -1.   FTR.....         26.25
    ZP15S203  ZP15..03            1.
    ZP15T103  BP15..T1           -1.   FTR.....           7.6
    ZP15T103  TYEAR3T1            1.   ZP15..03            1.
    ZP15T203  BP15..T2           -1.   FTR.....           8.7
    ZP15T203  TYEAR3T2            1.   ZP15..03            1.
    ZP15T303  BP15..T3           -1.   FTR.....          8.35
    ZP15T303  TYEAR3T3            1.   ZP15..03            1.
    ZP15T403  BP15..T4           -1.   FTR.....         10.45
    ZP15T403  TYEAR3T4            1.   ZP15..03            1.
    ZP25S103  BP25..S1           -1.   FTR.....          25.4
    ZP25S103  ZP25..03            1.
    ZP25S203  BP25..S2           -1.   FTR.....         26.25
    ZP25S203  ZP25..03            1.
    ZP25T103  BP25..T1           -1.   FTR.....           7.6
    ZP25T103  TYEAR3T1            1.   ZP25..03            1.
    ZP25T203  BP25..T2           -1.   FTR.....           8.7
    ZP25T203  TYEAR3T2            1.   ZP25..03            1.
    ZP25T303  BP25..T3           -1.   FTR.....          8.35
    ZP25T303  TYEAR3T3            1.   ZP25..03            1.
    ZP25T403  BP25..T4           -1.   FTR.....         10.45
    ZP25T403  TYEAR3T4            1.   ZP25..03            1.
    ZP16S103  BP16..S1           -1.   FTR.....          25.4
    ZP16S103  ZP16..03            1.
    ZP16S203  BP16..S2           -1.   FTR.....         26.25
    ZP16S203  ZP16..03            1.
    ZP16T103  BP16..T1           -1.   FTR.....           7.6
    ZP16T103  TYEAR3T1            1.   ZP16..03            1.
    ZP16T203  BP16..T2           -1.   FTR.....           8.7
    ZP16T203  TYEAR3T2            1.   ZP16..03            1.
    ZP16T303  BP16..T3           -1.   FTR.....          8.35
    ZP16T303  TYEAR3T3            1.   ZP16..03            1.
    ZP16T403  BP16..T4           -1.   FTR.....         10.45
    ZP16T403  TYEAR3T4            1.   ZP16..03            1.
    ZP26S103  BP26..S1           -1.   FTR.....          25.4
    ZP26S103  ZP26..03            1.
    ZP26S203  BP26..S2           -1.   FTR.....         26.25
    ZP26S203  ZP26..03            1.
    ZP26T103  BP26..T1           -1.   FTR.....           7.6
    ZP26T103  TYEAR3T1            1.   ZP26..03            1.
    ZP26T203  BP26..T2           -1.   FTR.....           8.7
    ZP26T203  TYEAR3T2            1.   ZP26..03            1.
    ZP26T303  BP26..T3           -1.   FTR.....          8.35
    ZP26T303  TYEAR3T3            1.   ZP26..03            1.
    ZP26T403  BP26..T4           -1.   FTR.....         10.45
    ZP26T403  TYEAR3T4            1.   ZP26..03            1.
    ZP11S104  BP11..S1           -1.   FTR.....          30.4
    ZP11S104  ZP11..04            1.
    ZP11S204  BP11..S2           -1.   FTR.....         30.35
    ZP11S204  ZP11..04            1.
    ZP11T104  BP11..T1           -1.   FTR.....            7.
    ZP11T104  TYEAR1T1            1.   ZP11..04            1.
    ZP11T204  BP11..T2           -1.   FTR.....          11.8
    ZP11T204  TYEAR1T2            1.   ZP11..04            1.
    ZP11T304  BP11..T3           -1.   FTR.....          10.4
    ZP11T304  TYEAR1T3            1.   ZP11..04            1.
    ZP11T404  BP11..T4           -1.   FTR.....          14.5
    ZP11T404  TYEAR1T4            1.   ZP11..04            1.
    ZP21S104  BP21..S1           -1.   FTR.....          30.4
    ZP21S104  ZP21..04            1.
    ZP21S204  BP21..S2           -1.   FTR.....         30.35
    ZP21S204  ZP21..04            1.
    ZP21T104  BP21..T1           -1.   FTR.....            7.
    ZP21T104  TYEAR1T1            1.   ZP21..04            1.
    ZP21T204  BP21..T2           -1.   FTR.....          11.8
    ZP21T204  TYEAR1T2            1.   ZP21..04            1.
    ZP21T304  BP21..T3           -1.   FTR.....          10.4
    ZP21T304  TYEAR1T3            1.   ZP21..04            1.
    ZP21T404  BP21..T4           -1.   FTR.....          14.5
    ZP21T404  TYEAR1T4            1.   ZP21..04            1.
    ZP12S104  BP12..S1           -1.   FTR.....          30.4
    ZP12S104  ZP12..04            1.
    ZP12S204  BP12..S2           -1.   FTR.....         30.35
    ZP12S204  ZP12..04            1.
    ZP12T104  BP12..T1           -1.   FTR.....            7.
    ZP12T104  TYEAR1T1            1.   ZP12..04            1.
    ZP12T204  BP12..T2           -1.   FTR.....          11.8
    ZP12T204  TYEAR1T2            1.   ZP12..04            1.
    ZP12T304  BP12..T3           -1.   FTR.....          10.4
    ZP12T304  TYEAR1T3            1.   ZP12..04            1.
    ZP12T404  BP12..T4           -1.   FTR.....          14.5
    ZP12T404  TYEAR1T4            1.   ZP12..04            1.
    ZP22S104  BP22..S1           -1.   FTR.....          30.4
    ZP22S104  ZP22..04            1.
    ZP22S204  BP22..S2           -1.   FTR.....         30.35
    ZP22S204  ZP22..04            1.
    ZP22T104  BP22..T1           -1.   FTR.....            7.
    ZP22T104  TYEAR1T1            1.   ZP22..04            1.
    ZP22T204  BP22..T2           -1.   FTR.....          11.8
    ZP22T204  TYEAR1T2            1.   ZP22..04            1.
    ZP22T304  BP22..T3           -1.   FTR.....          10.4
    ZP22T304  TYEAR1T3            1.   ZP22..04            1.
    ZP22T404  BP22..T4           -1.   FTR.....          14.5
    ZP22T404  TYEAR1T4            1.   ZP22..04            1.
    ZP13S104  BP13..S1           -1.   FTR.....          30.4
    ZP13S104  ZP13..04            1.
    ZP13S204  BP13..S2           -1.   FTR.....         30.35
    ZP13S204  ZP13..04            1.
    ZP13T104  BP13..T1           -1.   FTR.....            7.
    ZP13T104  TYEAR2T1            1.   ZP13..04            1.
    ZP13T204  BP13..T2           -1.   FTR.....          11.8
    ZP13T204  TYEAR2T2            1.   ZP13..04            1.
    ZP13T304  BP13..T3           -1.   FTR.....          10.4
    ZP13T304  TYEAR2T3            1.   ZP13..04            1.
    ZP13T404  BP13..T4           -1.   FTR.....          14.5
    ZP13T404  TYEAR2T4            1.   ZP13..04            1.
    ZP23S104  BP23..S1           -1.   FTR.....          30.4
    ZP23S104  ZP23..04            1.
    ZP23S204  BP23..S2           -1.   FTR.....         30.35
    ZP23S204  ZP23..04            1.
    ZP23T104  BP23..T1           -1.   FTR.....            7.
    ZP23T104  TYEAR2T1            1.   ZP23..04            1.
    ZP23T204  BP23..T2           -1.   FTR.....          11.8
    ZP23T204  TYEAR2T2            1.   ZP23..04            1.
    ZP23T304  BP23..T3           -1.   FTR.....          10.4
    ZP23T304  TYEAR2T3            1.   ZP23..04            1.
    ZP23T404  BP23..T4           -1.   FTR.....          14.5
    ZP23T404  TYEAR2T4            1.   ZP23..04            1.
    ZP14S104  BP14..S1           -1.   FTR.....          30.4
    ZP14S104  ZP14..04            1.
    ZP14S204  BP14..S2           -1.   FTR.....         30.35
    ZP14S204  ZP14..04            1.
    ZP14T104  BP14..T1           -1.   FTR.....            7.
    ZP14T104  TYEAR2T1            1.   ZP14..04            1.
    ZP14T204  BP14..T2           -1.   FTR.....          11.8
    ZP14T204  TYEAR2T2            1.   ZP14..04            1.
    ZP14T304  BP14..T3           -1.   FTR.....          10.4
    ZP14T304  TYEAR2T3            1.   ZP14..04            1.
    ZP14T404  BP14..T4           -1.   FTR.....          14.5
    ZP14T404  TYEAR2T4            1.   ZP14..04            1.
    ZP24S104  BP24..S1           -1.   FTR.....          30.4
    ZP24S104  ZP24..04            1.
    ZP24S204  BP24..S2           -1.   FTR.....         30.35
    ZP24S204  ZP24..04            1.
    ZP24T104  BP24..T1           -1.   FTR.....            7.
    ZP24T104  TYEAR2T1            1.   ZP24..04            1.
    ZP24T204  BP24..T2           -1.   FTR.....          11.8
    ZP24T204  TYEAR2T2            1.   ZP24..04            1.
    ZP24T304  BP24..T3           -1.   FTR.....          10.4
    ZP24T304  TYEAR2T3            1.   ZP24..04            1.
    ZP24T404  BP24..T4           -1.   FTR.....          14.5
    ZP24T404  TYEAR2T4            1.   ZP24..04            1.
    ZP15S104  BP15..S1           -1.   FTR.....          30.4
    ZP15S104  ZP15..04            1.
    ZP15S204  BP15..S2           -1.   FTR.....         30.35
    ZP15S204  ZP15..04            1.
    ZP15T104  BP15..T1           -1.   FTR.....            7.
    ZP15T104  TYEAR3T1            1.   ZP15..04            1.
    ZP15T204  BP15..T2           -1.   FTR.....          11.8
    ZP15T204  TYEAR3T2            1.   ZP15..04            1.
    ZP15T304  BP15..T3           -1.   FTR.....          10.4
    ZP15T304  TYEAR3T3            1.   ZP15..04            1.
    ZP15T404  BP15..T4           -1.   FTR.....          14.5
    ZP15T404  TYEAR3T4            1.   ZP15..04            1.
    ZP25S104  BP25..S1           -1.   FTR.....          30.4
    ZP25S104  ZP25..04            1.
    ZP25S204  BP25..S2           -1.   FTR.....         30.35
    ZP25S204  ZP25..04            1.
    ZP25T104  BP25..T1           -1.   FTR.....            7.
    ZP25T104  TYEAR3T1            1.   ZP25..04            1.
    ZP25T204  BP25..T2           -1.   FTR.....          11.8
    ZP25T204  TYEAR3T2            1.   ZP25..04            1.
    ZP25T304  BP25..T3           -1.   FTR.....          10.4
    ZP25T304  TYEAR3T3            1.   ZP25..04            1.
    ZP25T404  BP25..T4           -1.   FTR.....          14.5
    ZP25T404  TYEAR3T4            1.   ZP25..04            1.
    ZP16S104  BP16..S1           -1.   FTR.....          30.4
    ZP16S104  ZP16..04            1.
    ZP16S204  BP16..S2           -1.   FTR.....         30.35
    ZP16S204  ZP16..04            1.
    ZP16T104  BP16..T1           -1.   FTR.....            7.
    ZP16T104  TYEAR3T1            1.   ZP16..04            1.
    ZP16T204  BP16..T2           -1.   FTR.....          11.8
    ZP16T204  TYEAR3T2            1.   ZP16..04            1.
    ZP16T304  BP16..T3           -1.   FTR.....          10.4
    ZP16T304  TYEAR3T3            1.   ZP16..04            1.
    ZP16T404  BP16..T4           -1.   FTR.....          14.5
    ZP16T404  TYEAR3T4            1.   ZP16..04            1.
    ZP26S104  BP26..S1           -1.   FTR.....          30.4
    ZP26S104  ZP26..04            1.
    ZP26S204  BP26..S2           -1.   FTR.....         30.35
    ZP26S204  ZP26..04            1.
    ZP26T104  BP26..T1           -1.   FTR.....            7.
    ZP26T104  TYEAR3T1            1.   ZP26..04            1.
    ZP26T204  BP26..T2           -1.   FTR.....          11.8
    ZP26T204  TYEAR3T2            1.   ZP26..04            1.
    ZP26T304  BP26..T3           -1.   FTR.....          10.4
    ZP26T304  TYEAR3T3            1.   ZP26..04            1.
    ZP26T404  BP26..T4           -1.   FTR.....          14.5
    ZP26T404  TYEAR3T4            1.   ZP26..04            1.
    ZP11S105  BP11..S1           -1.   FTR.....           29.
    ZP11S105  ZP11..05            1.
    ZP11S205  BP11..S2           -1.   FTR.....          26.7
    ZP11S205  ZP11..05            1.
    ZP11T105  BP11..T1           -1.   FTR.....         11.25
    ZP11T105  TYEAR1T1            1.   ZP11..05            1.
    ZP11T205  BP11..T2           -1.   FTR.....           7.6
    ZP11T205  TYEAR1T2            1.   ZP11..05            1.
    ZP11T305  BP11..T3           -1.   FTR.....           6.1
    ZP11T305  TYEAR1T3            1.   ZP11..05            1.
    ZP11T405  BP11..T4           -1.   FTR.....          6.85
    ZP11T405  TYEAR1T4            1.   ZP11..05            1.
    ZP21S105  BP21..S1           -1.   FTR.....           29.
    ZP21S105  ZP21..05            1.
    ZP21S205  BP21..S2           -1.   FTR.....          26.7
    ZP21S205  ZP21..05            1.
    ZP21T105  BP21..T1           -1.   FTR.....         11.25
    ZP21T105  TYEAR1T1            1.   ZP21..05            1.
    ZP21T205  BP21..T2           -1.   FTR.....           7.6
    ZP21T205  TYEAR1T2            1.   ZP21..05            1.
    ZP21T305  BP21..T3           -1.   FTR.....           6.1
    ZP21T305  TYEAR1T3            1.   ZP21..05            1.
    ZP21T405  BP21..T4           -1.   FTR.....          6.85
    ZP21T405  TYEAR1T4            1.   ZP21..05            1.
    ZP12S105  BP12..S1           -1.   FTR.....           29.
    ZP12S105  ZP12..05            1.
    ZP12S205  BP12..S2           -1.   FTR.....          26.7
    ZP12S205  ZP12..05            1.
    ZP12T105  BP12..T1           -1.   FTR.....         11.25
    ZP12T105  TYEAR1T1            1.   ZP12..05            1.
    ZP12T205  BP12..T2           -1.   FTR.....           7.6
    ZP12T205  TYEAR1T2            1.   ZP12..05            1.
    ZP12T305  BP12..T3           -1.   FTR.....           6.1
    ZP12T305  TYEAR1T3            1.   ZP12..05            1.
    ZP12T405  BP12..T4           -1.   FTR.....          6.85
    ZP12T405  TYEAR1T4            1.   ZP12..05            1.
    ZP22S105  BP22..S1           -1.   FTR.....           29.
    ZP22S105  ZP22..05            1.
    ZP22S205  BP22..S2           -1.   FTR.....          26.7
    ZP22S205  ZP22..05            1.
    ZP22T105  BP22..T1           -1.   FTR.....         11.25
    ZP22T105  TYEAR1T1            1.   ZP22..05            1.
    ZP22T205  BP22..T2           -1.   FTR.....           7.6
    ZP22T205  TYEAR1T2            1.   ZP22..05            1.
    ZP22T305  BP22..T3           -1.   FTR.....           6.1
    ZP22T305  TYEAR1T3            1.   ZP22..05            1.
    ZP22T405  BP22..T4           -1.   FTR.....          6.85
    ZP22T405  TYEAR1T4            1.   ZP22..05            1.
    ZP13S105  BP13..S1           -1.   FTR.....           29.
    ZP13S105  ZP13..05            1.
    ZP13S205  BP13..S2           -1.   FTR.....          26.7
    ZP13S205  ZP13..05            1.
    ZP13T105  BP13..T1           -1.   FTR.....         11.25
    ZP13T105  TYEAR2T1            1.   ZP13..05            1.
    ZP13T205  BP13..T2           -1.   FTR.....           7.6
    ZP13T205  TYEAR2T2            1.   ZP13..05            1.
    ZP13T305  BP13..T3           -1.   FTR.....           6.1
    ZP13T305  TYEAR2T3            1.   ZP13..05            1.
    ZP13T405  BP13..T4           -1.   FTR.....          6.85
    ZP13T405  TYEAR2T4            1.   ZP13..05            1.
    ZP23S105  BP23..S1           -1.   FTR.....           29.
    ZP23S105  ZP23..05            1.
    ZP23S205  BP23..S2           -1.   FTR.....          26.7
    ZP23S205  ZP23..05            1.
    ZP23T105  BP23..T1           -1.   FTR.....         11.25
    ZP23T105  TYEAR2T1            1.   ZP23..05            1.
    ZP23T205  BP23..T2           -1.   FTR.....           7.6
    ZP23T205  TYEAR2T2            1.   ZP23..05            1.
    ZP23T305  BP23..T3           -1.   FTR.....           6.1
    ZP23T305  TYEAR2T3            1.   ZP23..05            1.
    ZP23T405  BP23..T4           -1.   FTR.....          6.85
    ZP23T405  TYEAR2T4            1.   ZP23..05            1.
    ZP14S105  BP14..S1           -1.   FTR.....           29.
    ZP14S105  ZP14..05            1.
    ZP14S205  BP14..S2           -1.   FTR.....          26.7
    ZP14S205  ZP14..05            1.
    ZP14T105  BP14..T1           -1.   FTR.....         11.25
    ZP14T105  TYEAR2T1            1.   ZP14..05            1.
    ZP14T205  BP14..T2           -1.   FTR.....           7.6
    ZP14T205  TYEAR2T2            1.   ZP14..05            1.
    ZP14T305  BP14..T3           -1.   FTR.....           6.1
    ZP14T305  TYEAR2T3            1.   ZP14..05            1.
    ZP14T405  BP14..T4           -1.   FTR.....          6.85
    ZP14T405  TYEAR2T4            1.   ZP14..05            1.
    ZP24S105  BP24..S1           -1.   FTR.....           29.
    ZP24S105  ZP24..05            1.
    ZP24S205  BP24..S2           -1.   FTR.....          26.7
    ZP24S205  ZP24..05            1.
    ZP24T105  BP24..T1           -1.   FTR.....         11.25
    ZP24T105  TYEAR2T1            1.   ZP24..05            1.
    ZP24T205  BP24..T2           -1.   FTR.....           7.6
    ZP24T205  TYEAR2T2            1.   ZP24..05            1.
    ZP24T305  BP24..T3           -1.   FTR.....           6.1
    ZP24T305  TYEAR2T3            1.   ZP24..05            1.
    ZP24T405  BP24..T4           -1.   FTR.....          6.85
    ZP24T405  TYEAR2T4            1.   ZP24..05            1.
    ZP15S105  BP15..S1           -1.   FTR.....           29.
    ZP15S105  ZP15..05            1.
    ZP15S205  BP15..S2           -1.   FTR.....          26.7
    ZP15S205  ZP15..05            1.
    ZP15T105  BP15..T1           -1.   FTR.....         11.25
    ZP15T105  TYEAR3T1            1.   ZP15..05            1.
    ZP15T205  BP15..T2           -1.   FTR.....           7.6
    ZP15T205  TYEAR3T2            1.   ZP15..05            1.
    ZP15T305  BP15..T3           -1.   FTR.....           6.1
    ZP15T305  TYEAR3T3            1.   ZP15..05            1.
    ZP15T405  BP15..T4           -1.   FTR.....          6.85
    ZP15T405  TYEAR3T4            1.   ZP15..05            1.
    ZP25S105  BP25..S1           -1.   FTR.....           29.
    ZP25S105  ZP25..05            1.
    ZP25S205  BP25..S2           -1.   FTR.....          26.7
    ZP25S205  ZP25..05            1.
    ZP25T105  BP25..T1           -1.   FTR.....         11.25
    ZP25T105  TYEAR3T1            1.   ZP25..05            1.
    ZP25T205  BP25..T2           -1.   FTR.....           7.6
    ZP25T205  TYEAR3T2            1.   ZP25..05            1.
    ZP25T305  BP25..T3           -1.   FTR.....           6.1
    ZP25T305  TYEAR3T3            1.   ZP25..05            1.
    ZP25T405  BP25..T4           -1.   FTR.....          6.85
    ZP25T405  TYEAR3T4            1.   ZP25..05            1.
    ZP16S105  BP16..S1           -1.   FTR.....           29.
    ZP16S105  ZP16..05            1.
    ZP16S205  BP16..S2           -1.   FTR.....          26.7
    ZP16S205  ZP16..05            1.
    ZP16T105  BP16..T1           -1.   FTR.....         11.25
    ZP16T105  TYEAR3T1            1.   ZP16..05            1.
    ZP16T205  BP16..T2           -1.   FTR.....           7.6
    ZP16T205  TYEAR3T2            1.   ZP16..05            1.
    ZP16T305  BP16..T3           -1.   FTR.....           6.1
    ZP16T305  TYEAR3T3            1.   ZP16..05            1.
    ZP16T405  BP16..T4           -1.   FTR.....          6.85
    ZP16T405  TYEAR3T4            1.   ZP16..05            1.
    ZP26S105  BP26..S1           -1.   FTR.....           29.
    ZP26S105  ZP26..05            1.
    ZP26S205  BP26..S2           -1.   FTR.....          26.7
    ZP26S205  ZP26..05            1.
    ZP26T105  BP26..T1           -1.   FTR.....         11.25
    ZP26T105  TYEAR3T1            1.   ZP26..05            1.
    ZP26T205  BP26..T2           -1.   FTR.....           7.6
    ZP26T205  TYEAR3T2            1.   ZP26..05            1.
    ZP26T305  BP26..T3           -1.   FTR.....           6.1
    ZP26T305  TYEAR3T3            1.   ZP26..05            1.
    ZP26T405  BP26..T4           -1.   FTR.....          6.85
    ZP26T405  TYEAR3T4            1.   ZP26..05            1.
    ZP11S106  BP11..S1           -1.   FTR.....           28.
    ZP11S106  ZP11..06            1.
    ZP11S206  BP11..S2           -1.   FTR.....          30.1
    ZP11S206  ZP11..06            1.
    ZP11T106  BP11..T1           -1.   FTR.....            8.
    ZP11T106  TYEAR1T1            1.   ZP11..06            1.
    ZP11T206  BP11..T2           -1.   FTR.....          5.75
    ZP11T206  TYEAR1T2            1.   ZP11..06            1.
    ZP11T306  BP11..T3           -1.   FTR.....           9.5
    ZP11T306  TYEAR1T3            1.   ZP11..06            1.
    ZP11T406  BP11..T4           -1.   FTR.....         11.75
    ZP11T406  TYEAR1T4            1.   ZP11..06            1.
    ZP21S106  BP21..S1           -1.   FTR.....           28.
    ZP21S106  ZP21..06            1.
    ZP21S206  BP21..S2           -1.   FTR.....          30.1
    ZP21S206  ZP21..06            1.
    ZP21T106  BP21..T1           -1.   FTR.....            8.
    ZP21T106  TYEAR1T1            1.   ZP21..06            1.
    ZP21T206  BP21..T2           -1.   FTR.....          5.75
    ZP21T206  TYEAR1T2            1.   ZP21..06            1.
    ZP21T306  BP21..T3           -1.   FTR.....           9.5
    ZP21T306  TYEAR1T3            1.   ZP21..06            1.
    ZP21T406  BP21..T4           -1.   FTR.....         11.75
    ZP21T406  TYEAR1T4            1.   ZP21..06            1.
    ZP12S106  BP12..S1           -1.   FTR.....           28.
    ZP12S106  ZP12..06            1.
    ZP12S206  BP12..S2           -1.   FTR.....          30.1
    ZP12S206  ZP12..06            1.
    ZP12T106  BP12..T1           -1.   FTR.....            8.
    ZP12T106  TYEAR1T1            1.   ZP12..06            1.
    ZP12T206  BP12..T2           -1.   FTR.....          5.75
    ZP12T206  TYEAR1T2            1.   ZP12..06            1.
    ZP12T306  BP12..T3           -1.   FTR.....           9.5
    ZP12T306  TYEAR1T3            1.   ZP12..06            1.
    ZP12T406  BP12..T4           -1.   FTR.....         11.75
    ZP12T406  TYEAR1T4            1.   ZP12..06            1.
    ZP22S106  BP22..S1           -1.   FTR.....           28.
    ZP22S106  ZP22..06            1.
    ZP22S206  BP22..S2           -1.   FTR.....          30.1
    ZP22S206  ZP22..06            1.
    ZP22T106  BP22..T1           -1.   FTR.....            8.
    ZP22T106  TYEAR1T1            1.   ZP22..06            1.
    ZP22T206  BP22..T2           -1.   FTR.....          5.75
    ZP22T206  TYEAR1T2            1.   ZP22..06            1.
    ZP22T306  BP22..T3           -1.   FTR.....           9.5
    ZP22T306  TYEAR1T3            1.   ZP22..06            1.
    ZP22T406  BP22..T4           -1.   FTR.....         11.75
    ZP22T406  TYEAR1T4            1.   ZP22..06            1.
    ZP13S106  BP13..S1           -1.   FTR.....           28.
    ZP13S106  ZP13..06            1.
    ZP13S206  BP13..S2           -1.   FTR.....          30.1
    ZP13S206  ZP13..06            1.
    ZP13T106  BP13..T1           -1.   FTR.....            8.
    ZP13T106  TYEAR2T1            1.   ZP13..06            1.
    ZP13T206  BP13..T2           -1.   FTR.....          5.75
    ZP13T206  TYEAR2T2            1.   ZP13..06            1.
    ZP13T306  BP13..T3           -1.   FTR.....           9.5
    ZP13T306  TYEAR2T3            1.   ZP13..06            1.
    ZP13T406  BP13..T4           -1.   FTR.....         11.75
    ZP13T406  TYEAR2T4            1.   ZP13..06            1.
    ZP23S106  BP23..S1           -1.   FTR.....           28.
    ZP23S106  ZP23..06            1.
    ZP23S206  BP23..S2           -1.   FTR.....          30.1
    ZP23S206  ZP23..06            1.
    ZP23T106  BP23..T1           -1.   FTR.....            8.
    ZP23T106  TYEAR2T1            1.   ZP23..06            1.
    ZP23T206  BP23..T2           -1.   FTR.....          5.75
    ZP23T206  TYEAR2T2            1.   ZP23..06            1.
    ZP23T306  BP23..T3           -1.   FTR.....           9.5
    ZP23T306  TYEAR2T3            1.   ZP23..06            1.
    ZP23T406  BP23..T4           -1.   FTR.....         11.75
    ZP23T406  TYEAR2T4            1.   ZP23..06            1.
    ZP14S106  BP14..S1           -1.   FTR.....           28.
    ZP14S106  ZP14..06            1.
    ZP14S206  BP14..S2           -1.   FTR.....          30.1
    ZP14S206  ZP14..06            1.
    ZP14T106  BP14..T1           -1.   FTR.....            8.
    ZP14T106  TYEAR2T1            1.   ZP14..06            1.
    ZP14T206  BP14..T2           -1.   FTR.....          5.75
    ZP14T206  TYEAR2T2            1.   ZP14..06            1.
    ZP14T306  BP14..T3           -1.   FTR.....           9.5
    ZP14T306  TYEAR2T3            1.   ZP14..06            1.
    ZP14T406  BP14..T4           -1.   FTR.....         11.75
    ZP14T406  TYEAR2T4            1.   ZP14..06            1.
    ZP24S106  BP24..S1           -1.   FTR.....           28.
    ZP24S106  ZP24..06            1.
    ZP24S206  BP24..S2           -1.   FTR.....          30.1
    ZP24S206  ZP24..06            1.
    ZP24T106  BP24..T1           -1.   FTR.....            8.
    ZP24T106  TYEAR2T1            1.   ZP24..06            1.
    ZP24T206  BP24..T2           -1.   FTR.....          5.75
    ZP24T206  TYEAR2T2            1.   ZP24..06            1.
    ZP24T306  BP24..T3           -1.   FTR.....           9.5
    ZP24T306  TYEAR2T3            1.   ZP24..06            1.
    ZP24T406  BP24..T4           -1.   FTR.....         11.75
    ZP24T406  TYEAR2T4            1.   ZP24..06            1.
    ZP15S106  BP15..S1           -1.   FTR.....           28.
    ZP15S106  ZP15..06            1.
    ZP15S206  BP15..S2           -1.   FTR.....          30.1
    ZP15S206  ZP15..06            1.
    ZP15T106  BP15..T1           -1.   FTR.....            8.
    ZP15T106  TYEAR3T1            1.   ZP15..06            1.
    ZP15T206  BP15..T2           -1.   FTR.....          5.75
    ZP15T206  TYEAR3T2            1.   ZP15..06            1.
    ZP15T306  BP15..T3            1.   FTR.....           9.5
    ZP15T306  TYEAR3T3            1.   ZP15..06            1.
    ZP15T406  BP15..T4           -1.   FTR.....         11.75
    ZP15T406  TYEAR3T4            1.   ZP15..06            1.
    ZP25S106  BP25..S1           -1.   FTR.....           28.
    ZP25S106  ZP25..06            1.
    ZP25S206  BP25..S2           -1.   FTR.....          30.1
    ZP25S206  ZP25..06            1.
    ZP25T106  BP25..T1           -1.   FTR.....            8.
    ZP25T106  TYEAR3T1            1.   ZP25..06            1.
    ZP25T206  BP25..T2           -1.   FTR.....          5.75
    ZP25T206  TYEAR3T2            1.   ZP25..06            1.
    ZP25T306  BP25..T3           -1.   FTR.....           9.5
    ZP25T306  TYEAR3T3            1.   ZP25..06            1.
    ZP25T406  BP25..T4           -1.   FTR.....         11.75
    ZP25T406  TYEAR3T4            1.   ZP25..06            1.
    ZP16S106  BP16..S1           -1.   FTR.....           28.
    ZP16S106  ZP16..06            1.
    ZP16S206  BP16..S2           -1.   FTR.....          30.1
    ZP16S206  ZP16..06            1.
    ZP16T106  BP16..T1           -1.   FTR.....            8.
    ZP16T106  TYEAR3T1            1.   ZP16..06            1.
    ZP16T206  BP16..T2           -1.   FTR.....          5.75
    ZP16T206  TYEAR3T2            1.   ZP16..06            1.
    ZP16T306  BP16..T3           -1.   FTR.....           9.5
    ZP16T306  TYEAR3T3            1.   ZP16..06            1.
    ZP16T406  BP16..T4           -1.   FTR.....         11.75
    ZP16T406  TYEAR3T4            1.   ZP16..06            1.
    ZP26S106  BP26..S1           -1.   FTR.....           28.
    ZP26S106  ZP26..06            1.
    ZP26S206  BP26..S2           -1.   FTR.....          30.1
    ZP26S206  ZP26..06            1.
    ZP26T106  BP26..T1           -1.   FTR.....            8.
    ZP26T106  TYEAR3T1            1.   ZP26..06            1.
    ZP26T206  BP26..T2           -1.   FTR.....          5.75
    ZP26T206  TYEAR3T2            1.   ZP26..06            1.
    ZP26T306  BP26..T3           -1.   FTR.....           9.5
    ZP26T306  TYEAR3T3            1.   ZP26..06            1.
    ZP26T406  BP26..T4           -1.   FTR.....         11.75
    ZP26T406  TYEAR3T4            1.   ZP26..06            1.
    ZP11S107  BP11..S1           -1.   FTR.....           26.
    ZP11S107  ZP11..07            1.
    ZP11S207  BP11..S2           -1.   FTR.....          26.4
    ZP11S207  ZP11..07            1.
    ZP11T107  BP11..T1           -1.   FTR.....           7.6
    ZP11T107  TYEAR1T1            1.   ZP11..07            1.
    ZP11T207  BP11..T2           -1.   FTR.....           8.7
    ZP11T207  TYEAR1T2            1.   ZP11..07            1.
    ZP11T307  BP11..T3           -1.   FTR.....          8.35
    ZP11T307  TYEAR1T3            1.   ZP11..07            1.
    ZP11T407  BP11..T4           -1.   FTR.....         10.45
    ZP11T407  TYEAR1T4            1.   ZP11..07            1.
    ZP21S107  BP21..S1           -1.   FTR.....           26.
    ZP21S107  ZP21..07            1.
    ZP21S207  BP21..S2           -1.   FTR.....          26.4
    ZP21S207  ZP21..07            1.
    ZP21T107  BP21..T1           -1.   FTR.....           7.6
    ZP21T107  TYEAR1T1            1.   ZP21..07            1.
    ZP21T207  BP21..T2           -1.   FTR.....           8.7
    ZP21T207  TYEAR1T2            1.   ZP21..07            1.
    ZP21T307  BP21..T3           -1.   FTR.....          8.35
    ZP21T307  TYEAR1T3            1.   ZP21..07            1.
    ZP21T407  BP21..T4           -1.   FTR.....         10.45
    ZP21T407  TYEAR1T4            1.   ZP21..07            1.
    ZP12S107  BP12..S1           -1.   FTR.....           26.
    ZP12S107  ZP12..07            1.
    ZP12S207  BP12..S2           -1.   FTR.....          26.4
    ZP12S207  ZP12..07            1.
    ZP12T107  BP12..T1           -1.   FTR.....           7.6
    ZP12T107  TYEAR1T1            1.   ZP12..07            1.
    ZP12T207  BP12..T2           -1.   FTR.....           8.7
    ZP12T207  TYEAR1T2            1.   ZP12..07            1.
    ZP12T307  BP12..T3           -1.   FTR.....          8.35
    ZP12T307  TYEAR1T3            1.   ZP12..07            1.
    ZP12T407  BP12..T4           -1.   FTR.....         10.45
    ZP12T407  TYEAR1T4            1.   ZP12..07            1.
    ZP22S107  BP22..S1           -1.   FTR.....           26.
    ZP22S107  ZP22..07            1.
    ZP22S207  BP22..S2           -1.   FTR.....          26.4
    ZP22S207  ZP22..07            1.
    ZP22T107  BP22..T1           -1.   FTR.....           7.6
    ZP22T107  TYEAR1T1            1.   ZP22..07            1.
    ZP22T207  BP22..T2           -1.   FTR.....           8.7
    ZP22T207  TYEAR1T2            1.   ZP22..07            1.
    ZP22T307  BP22..T3           -1.   FTR.....          8.35
    ZP22T307  TYEAR1T3            1.   ZP22..07            1.
    ZP22T407  BP22..T4           -1.   FTR.....         10.45
    ZP22T407  TYEAR1T4            1.   ZP22..07            1.
    ZP13S107  BP13..S1           -1.   FTR.....           26.
    ZP13S107  ZP13..07            1.
    ZP13S207  BP13..S2           -1.   FTR.....          26.4
    ZP13S207  ZP13..07            1.
    ZP13T107  BP13..T1           -1.   FTR.....           7.6
    ZP13T107  TYEAR2T1            1.   ZP13..07            1.
    ZP13T207  BP13..T2           -1.   FTR.....           8.7
    ZP13T207  TYEAR2T2            1.   ZP13..07            1.
    ZP13T307  BP13..T3           -1.   FTR.....          8.35
    ZP13T307  TYEAR2T3            1.   ZP13..07            1.
    ZP13T407  BP13..T4           -1.   FTR.....         10.45
    ZP13T407  TYEAR2T4            1.   ZP13..07            1.
    ZP23S107  BP23..S1           -1.   FTR.....           26.
    ZP23S107  ZP23..07            1.
    ZP23S207  BP23..S2           -1.   FTR.....          26.4
    ZP23S207  ZP23..07            1.
    ZP23T107  BP23..T1           -1.   FTR.....           7.6
    ZP23T107  TYEAR2T1            1.   ZP23..07            1.
    ZP23T207  BP23..T2           -1.   FTR.....           8.7
    ZP23T207  TYEAR2T2            1.   ZP23..07            1.
    ZP23T307  BP23..T3           -1.   FTR.....          8.35
    ZP23T307  TYEAR2T3            1.   ZP23..07            1.
    ZP23T407  BP23..T4           -1.   FTR.....         10.45
    ZP23T407  TYEAR2T4            1.   ZP23..07            1.
    ZP14S107  BP14..S1           -1.   FTR.....           26.
    ZP14S107  ZP14..07            1.
    ZP14S207  BP14..S2           -1.   FTR.....          26.4
    ZP14S207  ZP14..07            1.
    ZP14T107  BP14..T1           -1.   FTR.....           7.6
    ZP14T107  TYEAR2T1            1.   ZP14..07            1.
    ZP14T207  BP14..T2           -1.   FTR.....           8.7
    ZP14T207  TYEAR2T2            1.   ZP14..07            1.
    ZP14T307  BP14..T3           -1.   FTR.....          8.35
    ZP14T307  TYEAR2T3            1.   ZP14..07            1.
    ZP14T407  BP14..T4           -1.   FTR.....         10.45
    ZP14T407  TYEAR2T4            1.   ZP14..07            1.
    ZP24S107  BP24..S1           -1.   FTR.....           26.
    ZP24S107  ZP24..07            1.
    ZP24S207  BP24..S2           -1.   FTR.....          26.4
    ZP24S207  ZP24..07            1.
    ZP24T107  BP24..T1           -1.   FTR.....           7.6
    ZP24T107  TYEAR2T1            1.   ZP24..07            1.
    ZP24T207  BP24..T2           -1.   FTR.....           8.7
    ZP24T207  TYEAR2T2            1.   ZP24..07            1.
    ZP24T307  BP24..T3           -1.   FTR.....          8.35
    ZP24T307  TYEAR2T3            1.   ZP24..07            1.
    ZP24T407  BP24..T4           -1.   FTR.....         10.45
    ZP24T407  TYEAR2T4            1.   ZP24..07            1.
    ZP15S107  BP15..S1           -1.   FTR.....           26.
    ZP15S107  ZP15..07            1.
    ZP15S207  BP15..S2           -1.   FTR.....          26.4
    ZP15S207  ZP15..07            1.
    ZP15T107  BP15..T1           -1.   FTR.....           7.6
    ZP15T107  TYEAR3T1            1.   ZP15..07            1.
    ZP15T207  BP15..T2           -1.   FTR.....           8.7
    ZP15T207  TYEAR3T2            1.   ZP15..07            1.
    ZP15T307  BP15..T3           -1.   FTR.....          8.35
    ZP15T307  TYEAR3T3            1.   ZP15..07            1.
    ZP15T407  BP15..T4           -1.   FTR.....         10.45
    ZP15T407  TYEAR3T4            1.   ZP15..07            1.
    ZP25S107  BP25..S1           -1.   FTR.....           26.
    ZP25S107  ZP25..07            1.
    ZP25S207  BP25..S2           -1.   FTR.....          26.4
    ZP25S207  ZP25..07            1.
    ZP25T107  BP25..T1           -1.   FTR.....           7.6
    ZP25T107  TYEAR3T1            1.   ZP25..07            1.
    ZP25T207  BP25..T2           -1.   FTR.....           8.7
    ZP25T207  TYEAR3T2            1.   ZP25..07            1.
    ZP25T307  BP25..T3           -1.   FTR.....          8.35
    ZP25T307  TYEAR3T3            1.   ZP25..07            1.
    ZP25T407  BP25..T4           -1.   FTR.....         10.45
    ZP25T407  TYEAR3T4            1.   ZP25..07            1.
    ZP16S107  BP16..S1           -1.   FTR.....           26.
    ZP16S107  ZP16..07            1.
    ZP16S207  BP16..S2           -1.   FTR.....          26.4
    ZP16S207  ZP16..07            1.
    ZP16T107  BP16..T1           -1.   FTR.....           7.6
    ZP16T107  TYEAR3T1            1.   ZP16..07            1.
    ZP16T207  BP16..T2           -1.   FTR.....           8.7
    ZP16T207  TYEAR3T2            1.   ZP16..07            1.
    ZP16T307  BP16..T3           -1.   FTR.....          8.35
    ZP16T307  TYEAR3T3            1.   ZP16..07            1.
    ZP16T407  BP16..T4           -1.   FTR.....         10.45
    ZP16T407  TYEAR3T4            1.   ZP16..07            1.
    ZP26S107  BP26..S1           -1.   FTR.....           26.
    ZP26S107  ZP26..07            1.
    ZP26S207  BP26..S2           -1.   FTR.....          26.4
    ZP26S207  ZP26..07            1.
    ZP26T107  BP26..T1           -1.   FTR.....           7.6
    ZP26T107  TYEAR3T1            1.   ZP26..07            1.
    ZP26T207  BP26..T2           -1.   FTR.....           8.7
    ZP26T207  TYEAR3T2            1.   ZP26..07            1.
    ZP26T307  BP26..T3           -1.   FTR.....          8.35
    ZP26T307  TYEAR3T3            1.   ZP26..07            1.
    ZP26T407  BP26..T4           -1.   FTR.....         10.45
    ZP26T407  TYEAR3T4            1.   ZP26..07            1.
    ZP11S108  BP11..S1           -1.   FTR.....           28.
    ZP11S108  ZP11..08            1.
    ZP11S208  BP11..S2           -1.   FTR.....          28.2
    ZP11S208  ZP11..08            1.
    ZP11T108  BP11..T1           -1.   FTR.....           8.7
    ZP11T108  TYEAR1T1            1.   ZP11..08            1.
    ZP11T208  BP11..T2           -1.   FTR.....           9.7
    ZP11T208  TYEAR1T2            1.   ZP11..08            1.
    ZP11T308  BP11..T3           -1.   FTR.....           7.3
    ZP11T308  TYEAR1T3            1.   ZP11..08            1.
    ZP11T408  BP11..T4           -1.   FTR.....           8.7
    ZP11T408  TYEAR1T4            1.   ZP11..08            1.
    ZP21S108  BP21..S1           -1.   FTR.....           28.
    ZP21S108  ZP21..08            1.
    ZP21S208  BP21..S2           -1.   FTR.....          28.2
    ZP21S208  ZP21..08            1.
    ZP21T108  BP21..T1           -1.   FTR.....           8.7
    ZP21T108  TYEAR1T1            1.   ZP21..08            1.
    ZP21T208  BP21..T2           -1.   FTR.....           9.7
    ZP21T208  TYEAR1T2            1.   ZP21..08            1.
    ZP21T308  BP21..T3           -1.   FTR.....           7.3
    ZP21T308  TYEAR1T3            1.   ZP21..08            1.
    ZP21T408  BP21..T4           -1.   FTR.....           8.7
    ZP21T408  TYEAR1T4            1.   ZP21..08            1.
    ZP12S108  BP12..S1           -1.   FTR.....           28.
    ZP12S108  ZP12..08            1.
    ZP12S208  BP12..S2           -1.   FTR.....          28.2
    ZP12S208  ZP12..08            1.
    ZP12T108  BP12..T1           -1.   FTR.....           8.7
    ZP12T108  TYEAR1T1            1.   ZP12..08            1.
    ZP12T208  BP12..T2           -1.   FTR.....           9.7
    ZP12T208  TYEAR1T2            1.   ZP12..08            1.
    ZP12T308  BP12..T3           -1.   FTR.....           7.3
    ZP12T308  TYEAR1T3            1.   ZP12..08            1.
    ZP12T408  BP12..T4           -1.   FTR.....           8.7
    ZP12T408  TYEAR1T4            1.   ZP12..08            1.
    ZP22S108  BP22..S1           -1.   FTR.....           28.
    ZP22S108  ZP22..08            1.
    ZP22S208  BP22..S2           -1.   FTR.....          28.2
    ZP22S208  ZP22..08            1.
    ZP22T108  BP22..T1           -1.   FTR.....           8.7
    ZP22T108  TYEAR1T1            1.   ZP22..08            1.
    ZP22T208  BP22..T2           -1.   FTR.....           9.7
    ZP22T208  TYEAR1T2            1.   ZP22..08            1.
    ZP22T308  BP22..T3           -1.   FTR.....           7.3
    ZP22T308  TYEAR1T3            1.   ZP22..08            1.
    ZP22T408  BP22..T4           -1.   FTR.....           8.7
    ZP22T408  TYEAR1T4            1.   ZP22..08            1.
    ZP13S108  BP13..S1           -1.   FTR.....           28.
    ZP13S108  ZP13..08            1.
    ZP13S208  BP13..S2           -1.   FTR.....          28.2
    ZP13S208  ZP13..08            1.
    ZP13T108  BP13..T1           -1.   FTR.....           8.7
    ZP13T108  TYEAR2T1            1.   ZP13..08            1.
    ZP13T208  BP13..T2           -1.   FTR.....           9.7
    ZP13T208  TYEAR2T2            1.   ZP13..08            1.
    ZP13T308  BP13..T3           -1.   FTR.....           7.3
    ZP13T308  TYEAR2T3            1.   ZP13..08            1.
    ZP13T408  BP13..T4           -1.   FTR.....           8.7
    ZP13T408  TYEAR2T4            1.   ZP13..08            1.
    ZP23S108  BP23..S1           -1.   FTR.....           28.
    ZP23S108  ZP23..08            1.
    ZP23S208  BP23..S2           -1.   FTR.....          28.2
    ZP23S208  ZP23..08            1.
    ZP23T108  BP23..T1           -1.   FTR.....           8.7
    ZP23T108  TYEAR2T1            1.   ZP23..08            1.
    ZP23T208  BP23..T2           -1.   FTR.....           9.7
    ZP23T208  TYEAR2T2            1.   ZP23..08            1.
    ZP23T308  BP23..T3           -1.   FTR.....           7.3
    ZP23T308  TYEAR2T3            1.   ZP23..08            1.
    ZP23T408  BP23..T4           -1.   FTR.....           8.7
    ZP23T408  TYEAR2T4            1.   ZP23..08            1.
    ZP14S108  BP14..S1           -1.   FTR.....           28.
    ZP14S108  ZP14..08            1.
    ZP14S208  BP14..S2           -1.   FTR.....          28.2
    ZP14S208  ZP14..08            1.
    ZP14T108  BP14..T1           -1.   FTR.....           8.7
    ZP14T108  TYEAR2T1            1.   ZP14..08            1.
    ZP14T208  BP14..T2           -1.   FTR.....           9.7
    ZP14T208  TYEAR2T2            1.   ZP14..08            1.
    ZP14T308  BP14..T3           -1.   FTR.....           7.3
    ZP14T308  TYEAR2T3            1.   ZP14..08            1.
    ZP14T408  BP14..T4           -1.   FTR.....           8.7
    ZP14T408  TYEAR2T4            1.   ZP14..08            1.
    ZP24S108  BP24..S1           -1.   FTR.....           28.
    ZP24S108  ZP24..08            1.
    ZP24S208  BP24..S2           -1.   FTR.....          28.2
    ZP24S208  ZP24..08            1.
    ZP24T108  BP24..T1           -1.   FTR.....           8.7
    ZP24T108  TYEAR2T1            1.   ZP24..08            1.
    ZP24T208  BP24..T2           -1.   FTR.....           9.7
    ZP24T208  TYEAR2T2            1.   ZP24..08            1.
    ZP24T308  BP24..T3           -1.   FTR.....           7.3
    ZP24T308  TYEAR2T3            1.   ZP24..08            1.
    ZP24T408  BP24..T4           -1.   FTR.....           8.7
    ZP24T408  TYEAR2T4            1.   ZP24..08            1.
    ZP15S108  BP15..S1           -1.   FTR.....           28.
    ZP15S108  ZP15..08            1.
    ZP15S208  BP15..S2           -1.   FTR.....          28.2
    ZP15S208  ZP15..08            1.
    ZP15T108  BP15..T1           -1.   FTR.....           8.7
    ZP15T108  TYEAR3T1            1.   ZP15..08            1.
    ZP15T208  BP15..T2           -1.   FTR.....           9.7
    ZP15T208  TYEAR3T2            1.   ZP15..08            1.
    ZP15T308  BP15..T3           -1.   FTR.....           7.3
    ZP15T308  TYEAR3T3            1.   ZP15..08            1.
    ZP15T408  BP15..T4           -1.   FTR.....           8.7
    ZP15T408  TYEAR3T4            1.   ZP15..08            1.
    ZP25S108  BP25..S1           -1.   FTR.....           28.
    ZP25S108  ZP25..08            1.
    ZP25S208  BP25..S2           -1.   FTR.....          28.2
    ZP25S208  ZP25..08            1.
    ZP25T108  BP25..T1           -1.   FTR.....           8.7
    ZP25T108  TYEAR3T1            1.   ZP25..08            1.
    ZP25T208  BP25..T2           -1.   FTR.....           9.7
    ZP25T208  TYEAR3T2            1.   ZP25..08            1.
    ZP25T308  BP25..T3           -1.   FTR.....           7.3
    ZP25T308  TYEAR3T3            1.   ZP25..08            1.
    ZP25T408  BP25..T4           -1.   FTR.....           8.7
    ZP25T408  TYEAR3T4            1.   ZP25..08            1.
    ZP16S108  BP16..S1           -1.   FTR.....           28.
    ZP16S108  ZP16..08            1.
    ZP16S208  BP16..S2           -1.   FTR.....          28.2
    ZP16S208  ZP16..08            1.
    ZP16T108  BP16..T1           -1.   FTR.....           8.7
    ZP16T108  TYEAR3T1            1.   ZP16..08            1.
    ZP16T208  BP16..T2           -1.   FTR.....           9.7
    ZP16T208  TYEAR3T2            1.   ZP16..08            1.
    ZP16T308  BP16..T3           -1.   FTR.....           7.3
    ZP16T308  TYEAR3T3            1.   ZP16..08            1.
    ZP16T408  BP16..T4           -1.   FTR.....           8.7
    ZP16T408  TYEAR3T4            1.   ZP16..08            1.
    ZP26S108  BP26..S1           -1.   FTR.....           28.
    ZP26S108  ZP26..08            1.
    ZP26S208  BP26..S2           -1.   FTR.....          28.2
    ZP26S208  ZP26..08            1.
    ZP26T108  BP26..T1           -1.   FTR.....           8.7
    ZP26T108  TYEAR3T1            1.   ZP26..08            1.
    ZP26T208  BP26..T2           -1.   FTR.....           9.7
    ZP26T208  TYEAR3T2            1.   ZP26..08            1.
    ZP26T308  BP26..T3           -1.   FTR.....           7.3
    ZP26T308  TYEAR3T3            1.   ZP26..08            1.
    ZP26T408  BP26..T4           -1.   FTR.....           8.7
    ZP26T408  TYEAR3T4            1.   ZP26..08            1.
    ZP11S109  BP11..S1           -1.   FTR.....          26.2
    ZP11S109  ZP11..09            1.
    ZP11S209  BP11..S2           -1.   FTR.....           28.
    ZP11S209  ZP11..09            1.
    ZP11T109  BP11..T1           -1.   FTR.....           9.7
    ZP11T109  TYEAR1T1            1.   ZP11..09            1.
    ZP11T209  BP11..T2           -1.   FTR.....         10.75
    ZP11T209  TYEAR1T2            1.   ZP11..09            1.
    ZP11T309  BP11..T3           -1.   FTR.....           6.1
    ZP11T309  TYEAR1T3            1.   ZP11..09            1.
    ZP11T409  BP11..T4           -1.   FTR.....          9.25
    ZP11T409  TYEAR1T4            1.   ZP11..09            1.
    ZP21S109  BP21..S1           -1.   FTR.....          26.2
    ZP21S109  ZP21..09            1.
    ZP21S209  BP21..S2           -1.   FTR.....           28.
    ZP21S209  ZP21..09            1.
    ZP21T109  BP21..T1           -1.   FTR.....           9.7
    ZP21T109  TYEAR1T1            1.   ZP21..09            1.
    ZP21T209  BP21..T2           -1.   FTR.....         10.75
    ZP21T209  TYEAR1T2            1.   ZP21..09            1.
    ZP21T309  BP21..T3           -1.   FTR.....           6.1
    ZP21T309  TYEAR1T3            1.   ZP21..09            1.
    ZP21T409  BP21..T4           -1.   FTR.....          9.25
    ZP21T409  TYEAR1T4            1.   ZP21..09            1.
    ZP12S109  BP12..S1           -1.   FTR.....          26.2
    ZP12S109  ZP12..09            1.
    ZP12S209  BP12..S2           -1.   FTR.....           28.
    ZP12S209  ZP12..09            1.
    ZP12T109  BP12..T1           -1.   FTR.....           9.7
    ZP12T109  TYEAR1T1            1.   ZP12..09            1.
    ZP12T209  BP12..T2           -1.   FTR.....         10.75
    ZP12T209  TYEAR1T2            1.   ZP12..09            1.
    ZP12T309  BP12..T3           -1.   FTR.....           6.1
    ZP12T309  TYEAR1T3            1.   ZP12..09            1.
    ZP12T409  BP12..T4           -1.   FTR.....          9.25
    ZP12T409  TYEAR1T4            1.   ZP12..09            1.
    ZP22S109  BP22..S1           -1.   FTR.....          26.2
    ZP22S109  ZP22..09            1.
    ZP22S209  BP22..S2           -1.   FTR.....           28.
    ZP22S209  ZP22..09            1.
    ZP22T109  BP22..T1           -1.   FTR.....           9.7
    ZP22T109  TYEAR1T1            1.   ZP22..09            1.
    ZP22T209  BP22..T2           -1.   FTR.....         10.75
    ZP22T209  TYEAR1T2            1.   ZP22..09            1.
    ZP22T309  BP22..T3           -1.   FTR.....           6.1
    ZP22T309  TYEAR1T3            1.   ZP22..09            1.
    ZP22T409  BP22..T4           -1.   FTR.....          9.25
    ZP22T409  TYEAR1T4            1.   ZP22..09            1.
    ZP13S109  BP13..S1           -1.   FTR.....          26.2
    ZP13S109  ZP13..09            1.
    ZP13S209  BP13..S2           -1.   FTR.....           28.
    ZP13S209  ZP13..09            1.
    ZP13T109  BP13..T1           -1.   FTR.....           9.7
    ZP13T109  TYEAR2T1            1.   ZP13..09            1.
    ZP13T209  BP13..T2           -1.   FTR.....         10.75
    ZP13T209  TYEAR2T2            1.   ZP13..09            1.
    ZP13T309  BP13..T3           -1.   FTR.....           6.1
    ZP13T309  TYEAR2T3            1.   ZP13..09            1.
    ZP13T409  BP13..T4           -1.   FTR.....          9.25
    ZP13T409  TYEAR2T4            1.   ZP13..09            1.
    ZP23S109  BP23..S1           -1.   FTR.....          26.2
    ZP23S109  ZP23..09            1.
    ZP23S209  BP23..S2           -1.   FTR.....           28.
    ZP23S209  ZP23..09            1.
    ZP23T109  BP23..T1           -1.   FTR.....           9.7
    ZP23T109  TYEAR2T1            1.   ZP23..09            1.
    ZP23T209  BP23..T2           -1.   FTR.....         10.75
    ZP23T209  TYEAR2T2            1.   ZP23..09            1.
    ZP23T309  BP23..T3           -1.   FTR.....           6.1
    ZP23T309  TYEAR2T3            1.   ZP23..09            1.
    ZP23T409  BP23..T4           -1.   FTR.....          9.25
    ZP23T409  TYEAR2T4            1.   ZP23..09            1.
    ZP14S109  BP14..S1           -1.   FTR.....          26.2
    ZP14S109  ZP14..09            1.
    ZP14S209  BP14..S2           -1.   FTR.....           28.
    ZP14S209  ZP14..09            1.
    ZP14T109  BP14..T1           -1.   FTR.....           9.7
    ZP14T109  TYEAR2T1            1.   ZP14..09            1.
    ZP14T209  BP14..T2           -1.   FTR.....         10.75
    ZP14T209  TYEAR2T2            1.   ZP14..09            1.
    ZP14T309  BP14..T3           -1.   FTR.....           6.1
    ZP14T309  TYEAR2T3            1.   ZP14..09            1.
    ZP14T409  BP14..T4           -1.   FTR.....          9.25
    ZP14T409  TYEAR2T4            1.   ZP14..09            1.
    ZP24S109  BP24..S1           -1.   FTR.....          26.2
    ZP24S109  ZP24..09            1.
    ZP24S209  BP24..S2           -1.   FTR.....           28.
    ZP24S209  ZP24..09            1.
    ZP24T109  BP24..T1           -1.   FTR.....           9.7
    ZP24T109  TYEAR2T1            1.   ZP24..09            1.
    ZP24T209  BP24..T2           -1.   FTR.....         10.75
    ZP24T209  TYEAR2T2            1.   ZP24..09            1.
    ZP24T309  BP24..T3           -1.   FTR.....           6.1
    ZP24T309  TYEAR2T3            1.   ZP24..09            1.
    ZP24T409  BP24..T4           -1.   FTR.....          9.25
    ZP24T409  TYEAR2T4            1.   ZP24..09            1.
    ZP15S109  BP15..S1           -1.   FTR.....          26.2
    ZP15S109  ZP15..09            1.
    ZP15S209  BP15..S2           -1.   FTR.....           28.
    ZP15S209  ZP15..09            1.
    ZP15T109  BP15..T1           -1.   FTR.....           9.7
    ZP15T109  TYEAR3T1            1.   ZP15..09            1.
    ZP15T209  BP15..T2           -1.   FTR.....         10.75
    ZP15T209  TYEAR3T2            1.   ZP15..09            1.
    ZP15T309  BP15..T3           -1.   FTR.....           6.1
    ZP15T309  TYEAR3T3            1.   ZP15..09            1.
    ZP15T409  BP15..T4           -1.   FTR.....          9.25
    ZP15T409  TYEAR3T4            1.   ZP15..09            1.
    ZP25S109  BP25..S1           -1.   FTR.....          26.2
    ZP25S109  ZP25..09            1.
    ZP25S209  BP25..S2           -1.   FTR.....           28.
    ZP25S209  ZP25..09            1.
    ZP25T109  BP25..T1           -1.   FTR.....           9.7
    ZP25T109  TYEAR3T1            1.   ZP25..09            1.
    ZP25T209  BP25..T2           -1.   FTR.....         10.75
    ZP25T209  TYEAR3T2            1.   ZP25..09            1.
    ZP25T309  BP25..T3           -1.   FTR.....           6.1
    ZP25T309  TYEAR3T3            1.   ZP25..09            1.
    ZP25T409  BP25..T4           -1.   FTR.....          9.25
    ZP25T409  TYEAR3T4            1.   ZP25..09            1.
    ZP16S109  BP16..S1           -1.   FTR.....          26.2
    ZP16S109  ZP16..09            1.
    ZP16S209  BP16..S2           -1.   FTR.....           28.
    ZP16S209  ZP16..09            1.
    ZP16T109  BP16..T1           -1.   FTR.....           9.7
    ZP16T109  TYEAR3T1            1.   ZP16..09            1.
    ZP16T209  BP16..T2           -1.   FTR.....         10.75
    ZP16T209  TYEAR3T2            1.   ZP16..09            1.
    ZP16T309  BP16..T3           -1.   FTR.....           6.1
    ZP16T309  TYEAR3T3            1.   ZP16..09            1.
    ZP16T409  BP16..T4           -1.   FTR.....          9.25
    ZP16T409  TYEAR3T4            1.   ZP16..09            1.
    ZP26S109  BP26..S1           -1.   FTR.....          26.2
    ZP26S109  ZP26..09            1.
    ZP26S209  BP26..S2           -1.   FTR.....           28.
    ZP26S209  ZP26..09            1.
    ZP26T109  BP26..T1           -1.   FTR.....           9.7
    ZP26T109  TYEAR3T1            1.   ZP26..09            1.
    ZP26T209  BP26..T2           -1.   FTR.....         10.75
    ZP26T209  TYEAR3T2            1.   ZP26..09            1.
    ZP26T309  BP26..T3           -1.   FTR.....           6.1
    ZP26T309  TYEAR3T3            1.   ZP26..09            1.
    ZP26T409  BP26..T4           -1.   FTR.....          9.25
    ZP26T409  TYEAR3T4            1.   ZP26..09            1.
RHS
    RHS00001  LY....T1            1.   LY....T2            1.
    RHS00001  LY....T3            1.   LY....T4            1.
    RHS00001  FIN...1.         1500.   FIN...2.         1500.
    RHS00001  FIN...3.         1500.   ZP11..01           .36
    RHS00001  ZP21..01           .44   ZP12..01           .32
    RHS00001  ZP22..01            .5   ZP13..01           .41
    RHS00001  ZP23..01           .48   ZP14..01           .36
    RHS00001  ZP24..01           .55   ZP15..01           .43
    RHS00001  ZP25..01            .5   ZP16..01           .38
    RHS00001  ZP26..01           .57   ZP11..02           .82
    RHS00001  ZP21..02            .8   ZP12..02           .72
    RHS00001  ZP22..02           .88   ZP13..02           .92
    RHS00001  ZP23..02           .87   ZP14..02           .81
    RHS00001  ZP24..02           .96   ZP15..02           .97
    RHS00001  ZP25..02            .9   ZP16..02           .85
    RHS00001  ZP26..02            1.   ZP11..03           .34
    RHS00001  ZP21..03           .73   ZP12..03           .27
    RHS00001  ZP22..03           .47   ZP13..03           .36
    RHS00001  ZP23..03           .76   ZP14..03           .29
    RHS00001  ZP24..03           .49   ZP15..03           .38
    RHS00001  ZP25..03           .79   ZP16..03            .3
    RHS00001  ZP26..03           .51   ZP11..04            .9
    RHS00001  ZP21..04            1.   ZP12..04            .8
    RHS00001  ZP22..04          1.13   ZP13..04            1.
    RHS00001  ZP23..04          1.09   ZP14..04           .89
    RHS00001  ZP24..04          1.23   ZP15..04          1.06
    RHS00001  ZP25..04          1.13   ZP16..04           .94
    RHS00001  ZP26..04          1.18   ZP11..05           .37
    RHS00001  ZP21..05           .47   ZP12..05           .27
    RHS00001  ZP22..05           .29   ZP13..05            .4
    RHS00001  ZP23..05           .49   ZP14..05           .29
    RHS00001  ZP24..05            .3   ZP15..05           .42
    RHS00001  ZP25..05           .51   ZP16..05            .3
    RHS00001  ZP26..05           .31   ZP11..06           .34
    RHS00001  ZP21..06           .77   ZP12..06           .22
    RHS00001  ZP22..06           .49   ZP13..06           .37
    RHS00001  ZP23..06           .84   ZP14..06           .24
    RHS00001  ZP24..06           .53   ZP15..06            .4
    RHS00001  ZP25..06           .88   ZP16..06           .26
    RHS00001  ZP26..06           .55   ZP11..07           .34
    RHS00001  ZP21..07           .73   ZP12..07           .27
    RHS00001  ZP22..07           .47   ZP13..07           .38
    RHS00001  ZP23..07            .8   ZP14..07            .3
    RHS00001  ZP24..07           .51   ZP15..07            .4
    RHS00001  ZP25..07           .83   ZP16..07           .32
    RHS00001  ZP26..07           .53   ZP11..08            .4
    RHS00001  ZP21..08           .53   ZP12..08           .29
    RHS00001  ZP22..08           .32   ZP13..08           .45
    RHS00001  ZP23..08           .58   ZP14..08           .32
    RHS00001  ZP24..08           .35   ZP15..08           .47
    RHS00001  ZP25..08            .6   ZP16..08           .34
    RHS00001  ZP26..08           .37   ZP11..09           .33
    RHS00001  ZP21..09           .43   ZP12..09           .24
    RHS00001  ZP22..09           .27   ZP13..09           .36
    RHS00001  ZP23..09           .47   ZP14..09           .27
    RHS00001  ZP24..09           .29   ZP15..09           .38
    RHS00001  ZP25..09           .49   ZP16..09           .28
    RHS00001  ZP26..09            .3
BOUNDS
 UP BOUND     A..1X1S1          100.
 UP BOUND     A..1X2S1           50.
 UP BOUND     A..1X3S1           50.
 UP BOUND     A..2X1S1           13.
 UP BOUND     A..2X2S1           10.
 UP BOUND     A..2X3S1          177.
 UP BOUND     A..3X1S1          100.
 UP BOUND     A..3X2S1           50.
 UP BOUND     A..3X3S1           50.
 UP BOUND     A..4X1S1           13.
 UP BOUND     A..4X2S1           10.
 UP BOUND     A..4X3S1          177.
 UP BOUND     A..5X1S1          100.
 UP BOUND     A..5X2S1           50.
 UP BOUND     A..5X3S1           50.
 UP BOUND     A..6X1S1           13.
 UP BOUND     A..6X2S1           10.
 UP BOUND     A..6X3S1          177.
 UP BOUND     A..1X1S2          100.
 UP BOUND     A..1X2S2           50.
 UP BOUND     A..1X3S2           50.
 UP BOUND     A..2X1S2           13.
 UP BOUND     A..2X2S2           10.
 UP BOUND     A..2X3S2          177.
 UP BOUND     A..3X1S2          100.
 UP BOUND     A..3X2S2           50.
 UP BOUND     A..3X3S2           50.
 UP BOUND     A..4X1S2           13.
 UP BOUND     A..4X2S2           10.
 UP BOUND     A..4X3S2          177.
 UP BOUND     A..5X1S2          100.
 UP BOUND     A..5X2S2           50.
 UP BOUND     A..5X3S2           50.
 UP BOUND     A..6X1S2           13.
 UP BOUND     A..6X2S2           10.
 UP BOUND     A..6X3S2          177.
 UP BOUND     C..1..T1           10.
 UP BOUND     C..2..T1           10.
 UP BOUND     C..3..T1           10.
 UP BOUND     C..4..T1           15.
 UP BOUND     C..5..T1           15.
 UP BOUND     C..6..T1           15.
 UP BOUND     C..1..T3           10.
 UP BOUND     C..2..T3           10.
 UP BOUND     C..3..T3           10.
 UP BOUND     C..4..T3           15.
 UP BOUND     C..5..T3           15.
 UP BOUND     C..6..T3           15.
 FX BOUND     KBU1S1T1            1.
 UP BOUND     KBU1S2T1            1.
 FX BOUND     KIN1..T1            1.
 UP BOUND     KVA1..T1            1.
 FX BOUND     KBU2S1T1            1.
 UP BOUND     KBU2S2T1            1.
 FX BOUND     KBU3S1T1            1.
 UP BOUND     KBU3S2T1            1.
 UP BOUND     KIN3..T1            1.
 UP BOUND     KVA3..T1            1.
 FX BOUND     KBU4S1T1            1.
 UP BOUND     KBU4S2T1            1.
 FX BOUND     KBU5S1T1            1.
 UP BOUND     KBU5S2T1            1.
 UP BOUND     KIN5..T1            1.
 FX BOUND     KVA5..T1            1.
 FX BOUND     KBU6S1T1            1.
 UP BOUND     KBU6S2T1            1.
 UP BOUND     KBU1S1T2            1.
 UP BOUND     KBU1S2T2            1.
 UP BOUND     KIN1..T2            1.
 UP BOUND     KVA1..T2            1.
 UP BOUND     KBU2S1T2            1.
 UP BOUND     KBU2S2T2            1.
 UP BOUND     KBU3S1T2            1.
 UP BOUND     KBU3S2T2            1.
 UP BOUND     KIN3..T2            1.
 UP BOUND     KVA3..T2            1.
 UP BOUND     KBU4S1T2            1.
 UP BOUND     KBU4S2T2            1.
 UP BOUND     KBU5S1T2            1.
 UP BOUND     KBU5S2T2            1.
 UP BOUND     KIN5..T2            1.
 UP BOUND     KVA5..T2            1.
 UP BOUND     KBU6S1T2            1.
 UP BOUND     KBU6S2T2            1.
 FX BOUND     KBU1S1T3            0.
 FX BOUND     KBU1S2T3            0.
 FX BOUND     KIN1..T3            1.
 UP BOUND     KVA1..T3            1.
 FX BOUND     KBU2S1T3            0.
 UP BOUND     KBU2S2T3            1.
 UP BOUND     KBU3S1T3            1.
 UP BOUND     KBU3S2T3            1.
 UP BOUND     KIN3..T3            1.
 FX BOUND     KVA3..T3            1.
 UP BOUND     KBU4S1T3            1.
 UP BOUND     KBU4S2T3            1.
 FX BOUND     KBU5S1T3            0.
 UP BOUND     KBU5S2T3            1.
 UP BOUND     KIN5..T3            1.
 FX BOUND     KVA5..T3            1.
 FX BOUND     KBU6S1T3            0.
 UP BOUND     KBU6S2T3            1.
 UP BOUND     KBU1S1T4            1.
 UP BOUND     KBU1S2T4            1.
 UP BOUND     KIN1..T4            1.
 UP BOUND     KVA1..T4            1.
 UP BOUND     KBU2S1T4            1.
 UP BOUND     KBU2S2T4            1.
 UP BOUND     KBU3S1T4            1.
 UP BOUND     KBU3S2T4            1.
 UP BOUND     KIN3..T4            1.
 UP BOUND     KVA3..T4            1.
 UP BOUND     KBU4S1T4            1.
 UP BOUND     KBU4S2T4            1.
 UP BOUND     KBU5S1T4            1.
 UP BOUND     KBU5S2T4            1.
 UP BOUND     KIN5..T4            1.
 UP BOUND     KVA5..T4            1.
 UP BOUND     KBU6S1T4            1.
 UP BOUND     KBU6S2T4            1.
ENDATA
