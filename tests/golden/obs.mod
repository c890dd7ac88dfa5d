eta_0                :14.
magnetic field [G]   :1200.
LOS velocity [km/s]  :0.5
Doppler width [A]    :0.07
damping              :0.05
gamma [deg]          :130.
phi   [deg]          :25.
S_0                  :0.2
S_1                  :0.75
v_mac [km/s]         :0
filling factor       :0.85

