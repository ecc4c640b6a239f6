eta_0:          40
magnetic field: 1000
LOS velocity:   0.2
Doppler width:  0.03
damping:        0.05
gamma:          150
phi:            110
S_0:            0.35
S_1:            0.5
v_mac:          0
filling factor: 1