function [Im]= Born_adj_MNAD(dx,dz,nt,dt,v,xr,zr,data,d_pml,ss,Auu,Auv,Auw,Avu,Avv,Avw,Awu,Awv,Aww)
% BORN_ADJ_MNAD  Drop-in replacement of SRC/migration/Born_adj_MNAD.m that runs
% the receiver back-propagation and the imaging condition on the GPU through
% mnad_mex (libmnad, MI355X).
%
% Same arguments and output as the reference (Im: interior column vector,
% already multiplied by T).  U_r of the reference (undefined there) is read
% as U; data may be the matrix the driver passes or the cell {datau, ...}.
% The source-side snapshots come from the fwd_PDE_exact_MNAD call with
% indx == 1 for shot ss: from GPU memory if it ran in this process, else
% from ./Wavefields/UU_<ss>.mnad.

if iscell(data)
    data = data{1} ;
end
try
    Im = mnad_mex('adj', dx, dz, nt, dt, v, xr, zr, data, d_pml, ss) ;
catch err
    f = ['./Wavefields/UU_' num2str(ss) '.mnad'] ;
    if ~strcmp(err.identifier, 'mnad:state') || ~exist(f, 'file')
        rethrow(err)
    end
    mnad_mex('load', ss, f) ;
    Im = mnad_mex('adj', dx, dz, nt, dt, v, xr, zr, data, d_pml, ss) ;
end
end
