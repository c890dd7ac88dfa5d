eta_0                :20
magnetic field [G]   :1100
LOS velocity [km/s]  :0.2
Doppler width [A]    :0.03
damping              :0.05
gamma [deg]          :120
phi   [deg]          :150
S_0                  :0.35
S_1                  :0.5
v_mac [km/s]         :0
filling factor       :1
