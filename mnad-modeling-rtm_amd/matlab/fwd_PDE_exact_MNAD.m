function data = fwd_PDE_exact_MNAD(dx,xs,xr,dz,zs,zr,v,dt,nt,w,d_pml,ss,Auu,Auv,Auw,Avu,Avv,Avw,Awu,Awv,Aww,indx)
% FWD_PDE_EXACT_MNAD  Drop-in replacement of SRC/MNAD/fwd_PDE_exact_MNAD.m that
% runs the shot on the GPU through mnad_mex (libmnad, MI355X).
%
% Same arguments and output as the reference: data = {datau}, [nt x nr].
% The nine operator matrices are accepted and not used (the kernel evaluates
% their entries from v, dx, dz, dt).  With indx == 1 the source wavefield
% stays on the GPU under the key ss for Born_adj_MNAD, and the illumination
% goes to ./Illuminations/illum_<ss>.mat as in the reference
% (main_Acoustic_RTM.m:239).  Only when the adjoint of the shot runs in
% another process (another parfor worker) does the source wavefield have to
% go to disk: set MNAD_SPILL=1 in the environment and it is also written to
% ./Wavefields/UU_<ss>.mnad, which Born_adj_MNAD loads when the shot is not
% on the GPU of its own process.

if nargin < 22 || isempty(indx)
    indx = 0 ;
end

if indx == 1
    [datau, illum] = mnad_mex('fwd', dx, dz, dt, d_pml, v, w, nt, xs, zs, xr, zr, ss, indx) ;
    if ~exist('Illuminations','dir'), mkdir Illuminations; end
    save(['./Illuminations/illum_' num2str(ss)],'illum')
    if strcmp(getenv('MNAD_SPILL'), '1')
        if ~exist('Wavefields','dir'), mkdir Wavefields; end
        mnad_mex('save', ss, ['./Wavefields/UU_' num2str(ss) '.mnad']) ;
    end
else
    datau = mnad_mex('fwd', dx, dz, dt, d_pml, v, w, nt, xs, zs, xr, zr, ss, indx) ;
end

data{1} = datau ;
end
