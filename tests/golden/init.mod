eta_0:          20
magnetic field: 1100
LOS velocity:   0.2
Doppler width:  0.03
damping:        0.05
gamma:          120
phi:            150
S_0:            0.35
S_1:            0.5
v_mac:          0
filling factor: 1